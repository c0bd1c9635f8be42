// Kernels of the on-device LP solver that takes the reduced LP from presolve
// (restarted primal-dual hybrid gradient, PDHG, on the Ruiz-scaled LP):
// the two half steps of an iteration, Ruiz scaling, sparse products for the
// norm estimate, and the residual sums of the KKT check.
//
// Every kernel that walks the rows of a CSR matrix (A, or AT for the
// columns) runs over row blocks built once per solve: either one row,
// reduced by the whole workgroup (every row of more than
// GPK_PDHG_LONG_ROW entries gets such a block), or up to 256 consecutive
// short rows holding at most GPK_PDHG_ENTRY_BUDGET entries, whose products
// are staged in LDS by coalesced loads and then summed one row per thread.
// A column of AT with thousands of entries therefore never runs on a single
// lane.  Each row's sum has a fixed order (given by the blocks, which depend
// only on the matrix), and every sum over a vector goes through per-block
// partial sums reduced in block order by one workgroup, so a solve is
// deterministic run to run.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "gp_device.hpp"
#include "gp_kernels.h"

using namespace gpdev;

namespace {

constexpr int kEntryBudget = GPK_PDHG_ENTRY_BUDGET;
constexpr int kStage = kEntryBudget / kBlock;  // staged entries per thread
constexpr int kUnroll = 4;                     // chains per thread on a long row
constexpr int kPartials = 4;
static_assert(kEntryBudget % kBlock == 0, "entry budget in whole strides");

__device__ inline double clampd(double v, double lo, double hi) { return fmin(fmax(v, lo), hi); }

struct SumOp {
  static __device__ double identity() { return 0.0; }
  static __device__ double combine(double a, double b) { return a + b; }
};
struct MaxOp {
  static __device__ double identity() { return 0.0; }
  static __device__ double combine(double a, double b) { return fmax(a, b); }
};

// Combination of v over the block in a fixed order (lanes by shuffle tree,
// waves in index order); every thread gets the result.
template <class C>
__device__ inline double block_combine(double v, double* s_w) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = C::combine(v, __shfl_down(v, d, 64));
  if (lane_id() == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = s_w[0];
  for (int w = 1; w < kBlock / 64; ++w) t = C::combine(t, s_w[w]);
  __syncthreads();
  return t;
}

// ---- row operations --------------------------------------------------------
// elem(p, col, val): the term entry p contributes; row(r, acc, t): consumes
// the row's combined terms (and, with kTerms, adds its KKT terms to t).

// y = M v.
struct OpSpmv {
  using Comb = SumOp;
  static constexpr bool kTerms = false;
  const double* v;
  double* out;
  __device__ double elem(int32_t, int32_t col, double val) const { return val * v[col]; }
  __device__ void row(int32_t r, double acc, double*) const { out[r] = acc; }
};

// Ruiz: d[r] /= sqrt(max |entry| of row r) (rows without a nonzero keep d).
struct OpRuiz {
  using Comb = MaxOp;
  static constexpr bool kTerms = false;
  double* d;
  __device__ double elem(int32_t, int32_t, double val) const { return fabs(val); }
  __device__ void row(int32_t r, double acc, double*) const {
    if (acc > 0.0) d[r] /= sqrt(acc);
  }
};

// Halpern weight of iteration k since the restart: (k + 1) / (k + 2), with
// k = step[2] (the count at the start of the launch) + j.
__device__ inline double halpern_weight(const double* step, int32_t j) {
  const double k = step[2] + static_cast<double>(j);
  return (k + 1.0) / (k + 2.0);
}

// Primal half step over the rows of AT: xn = proj_[l,u](x - tau (c - A^T y))
// and xbar = 2 xn - x (the extrapolated point the dual step reads).
// Averaged PDHG: x <- xn, x_acc += xn (the running sum).  Reflected Halpern
// PDHG: x_acc <- xn (the PDHG point T(z)), x <- w xbar + (1 - w) x0, since
// xbar is also the reflected point 2 T(z) - z.
template <bool kHalpern>
struct OpPrimal {
  using Comb = SumOp;
  static constexpr bool kTerms = false;
  const double* y;
  const double* c;
  const double* l;
  const double* u;
  const double* step;
  double* x;
  double* xbar;
  double* x_acc;
  const double* x0;
  int32_t j;
  __device__ double elem(int32_t, int32_t col, double val) const { return val * y[col]; }
  __device__ void row(int32_t r, double aty, double*) const {
    const double g = c[r] - aty;
    const double xo = x[r];
    const double xn = clampd(xo - step[0] * g, l[r], u[r]);
    const double xb = 2.0 * xn - xo;
    xbar[r] = xb;
    if (kHalpern) {
      const double w = halpern_weight(step, j);
      x_acc[r] = xn;
      x[r] = w * xb + (1.0 - w) * x0[r];
    } else {
      x[r] = xn;
      x_acc[r] += xn;
    }
  }
};

// Dual half step over the rows of A: yn = yt + sigma proj_[AL,AU](-yt / sigma)
// with yt = y - sigma A xbar (the Moreau form of the projected dual ascent).
// Averaged: y <- yn, y_acc += yn.  Reflected Halpern: y_acc <- yn,
// y <- w (2 yn - y) + (1 - w) y0.
template <bool kHalpern>
struct OpDual {
  using Comb = SumOp;
  static constexpr bool kTerms = false;
  const double* xbar;
  const double* AL;
  const double* AU;
  const double* step;
  double* y;
  double* y_acc;
  const double* y0;
  int32_t j;
  __device__ double elem(int32_t, int32_t col, double val) const { return val * xbar[col]; }
  __device__ void row(int32_t i, double ax, double*) const {
    const double sigma = step[1];
    const double yo = y[i];
    const double yt = yo - sigma * ax;
    const double yn = yt + sigma * clampd(-yt / sigma, AL[i], AU[i]);
    if (kHalpern) {
      const double w = halpern_weight(step, j);
      y_acc[i] = yn;
      y[i] = w * (2.0 * yn - yo) + (1.0 - w) * y0[i];
    } else {
      y[i] = yn;
      y_acc[i] += yn;
    }
  }
};

// KKT terms of the rows (A, v = x): [0] squared violation of Ax against
// [AL, AU], [1] the rows' dual-objective terms of y, [2] squared bound
// magnitude (the larger finite |AL|, |AU|), [3] y^2 where the bound of y's
// sign is infinite (a dual infeasibility).
struct OpRowTerms {
  using Comb = SumOp;
  static constexpr bool kTerms = true;
  const double* x;
  const double* AL;
  const double* AU;
  const double* y;
  __device__ double elem(int32_t, int32_t col, double val) const { return val * x[col]; }
  __device__ void row(int32_t i, double ax, double* t) const {
    const double lo = AL[i], hi = AU[i];
    const double viol = ax - clampd(ax, lo, hi);
    t[0] += viol * viol;
    const double yi = y[i];
    if (yi > 0.0) {
      if (isfinite(lo)) t[1] += yi * lo;
      else t[3] += yi * yi;
    } else if (yi < 0.0) {
      if (isfinite(hi)) t[1] += yi * hi;
      else t[3] += yi * yi;
    }
    const double b = fmax(isfinite(lo) ? fabs(lo) : 0.0, isfinite(hi) ? fabs(hi) : 0.0);
    t[2] += b * b;
  }
};

// KKT terms of the columns (AT, v = y): g = c - A^T y and lambda its
// projection on the reduced-cost cone of the column's bounds (lambda >= 0
// needs a finite lower bound, lambda <= 0 a finite upper one); [0] squared
// dual residual (g - lambda), [1] the columns' dual-objective terms, [2] c^2,
// [3] c x; z = lambda when z is given.
struct OpColTerms {
  using Comb = SumOp;
  static constexpr bool kTerms = true;
  const double* y;
  const double* c;
  const double* l;
  const double* u;
  const double* x;
  double* z;
  __device__ double elem(int32_t, int32_t col, double val) const { return val * y[col]; }
  __device__ void row(int32_t j, double aty, double* t) const {
    const double g = c[j] - aty;
    const bool lo_fin = isfinite(l[j]), hi_fin = isfinite(u[j]);
    const double lam = clampd(g, hi_fin ? -INFINITY : 0.0, lo_fin ? INFINITY : 0.0);
    if (z) z[j] = lam;
    t[0] += (g - lam) * (g - lam);
    if (lam > 0.0) t[1] += lam * l[j];
    else if (lam < 0.0) t[1] += lam * u[j];
    t[2] += c[j] * c[j];
    t[3] += c[j] * x[j];
  }
};

template <class Op>
__global__ __launch_bounds__(kBlock) void k_rows(gpk_csr M, const int32_t* __restrict__ blk, Op op,
                                                 double* __restrict__ partial) {
  using C = typename Op::Comb;
  __shared__ double s_val[kEntryBudget];
  __shared__ int32_t s_ptr[kBlock + 1];
  __shared__ double s_w[kBlock / 64];
  const int32_t r0 = blk[blockIdx.x], r1 = blk[blockIdx.x + 1];
  const int32_t nr = r1 - r0;
  const int tid = threadIdx.x;
  double t[kPartials] = {0.0, 0.0, 0.0, 0.0};
  if (nr == 1) {
    // one (long) row: kUnroll independent strided chains per thread (their
    // loads in flight together), combined in a fixed order, then the tree
    const int32_t b = M.ptr[r0], e = M.ptr[r0 + 1];
    double acc[kUnroll];
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) acc[k] = C::identity();
    int32_t p = b + tid;
    for (; p + (kUnroll - 1) * kBlock < e; p += kUnroll * kBlock) {
      int32_t ci[kUnroll];
      double cv[kUnroll];
#pragma unroll
      for (int k = 0; k < kUnroll; ++k) {
        ci[k] = M.idx[p + k * kBlock];
        cv[k] = M.val[p + k * kBlock];
      }
#pragma unroll
      for (int k = 0; k < kUnroll; ++k) acc[k] = C::combine(acc[k], op.elem(p + k * kBlock, ci[k], cv[k]));
    }
#pragma unroll
    for (int k = 0; k < kUnroll - 1; ++k) {
      const int32_t q = p + k * kBlock;
      if (q < e) acc[k] = C::combine(acc[k], op.elem(q, M.idx[q], M.val[q]));
    }
    double a = acc[0];
#pragma unroll
    for (int k = 1; k < kUnroll; ++k) a = C::combine(a, acc[k]);
    a = block_combine<C>(a, s_w);
    if (tid == 0) op.row(r0, a, t);
  } else if (nr > 1) {
    const int32_t e0 = M.ptr[r0];
    const int32_t cnt = M.ptr[r1] - e0;
    if (nr <= kBlock && cnt <= kEntryBudget) {
      if (tid < nr) s_ptr[tid] = M.ptr[r0 + tid] - e0;
      if (tid == 0) s_ptr[nr] = cnt;
      // every load of the thread's (up to kStage) entries issued before the
      // gathers, every gather before the LDS stores
      int32_t ci[kStage];
      double cv[kStage];
#pragma unroll
      for (int k = 0; k < kStage; ++k) {
        const int32_t q = tid + k * kBlock;
        if (q < cnt) {
          ci[k] = M.idx[e0 + q];
          cv[k] = M.val[e0 + q];
        }
      }
#pragma unroll
      for (int k = 0; k < kStage; ++k) {
        const int32_t q = tid + k * kBlock;
        if (q < cnt) cv[k] = op.elem(e0 + q, ci[k], cv[k]);
      }
#pragma unroll
      for (int k = 0; k < kStage; ++k) {
        const int32_t q = tid + k * kBlock;
        if (q < cnt) s_val[q] = cv[k];
      }
      __syncthreads();
      if (tid < nr) {
        double acc = C::identity();
        for (int32_t q = s_ptr[tid]; q < s_ptr[tid + 1]; ++q) acc = C::combine(acc, s_val[q]);
        op.row(r0 + tid, acc, t);
      }
    } else {
      // not a block the host builds (more rows or entries than the block
      // stages): still correct, one row per thread from global memory
      for (int32_t r = r0 + tid; r < r1; r += kBlock) {
        double acc = C::identity();
        for (int32_t p = M.ptr[r]; p < M.ptr[r + 1]; ++p) acc = C::combine(acc, op.elem(p, M.idx[p], M.val[p]));
        op.row(r, acc, t);
      }
    }
  }
  if constexpr (Op::kTerms) {
    for (int k = 0; k < kPartials; ++k) {
      const double v = block_combine<SumOp>(t[k], s_w);
      if (tid == 0) partial[static_cast<int64_t>(blockIdx.x) * kPartials + k] = v;
    }
  }
}

// out[p] = M.val[p] * drow[row(p)] * dcol[idx[p]], over the same row blocks
// (entries strided over the block; the row of an entry of a multi-row block
// by binary search in the block's row pointers).
__global__ __launch_bounds__(kBlock) void k_rows_scale(gpk_csr M, const int32_t* __restrict__ blk,
                                                       const double* __restrict__ drow,
                                                       const double* __restrict__ dcol, double* __restrict__ out) {
  __shared__ int32_t s_ptr[kBlock + 1];
  const int32_t r0 = blk[blockIdx.x], r1 = blk[blockIdx.x + 1];
  const int32_t nr = r1 - r0;
  const int tid = threadIdx.x;
  if (nr == 1) {
    const double dr = drow[r0];
    for (int32_t p = M.ptr[r0] + tid; p < M.ptr[r0 + 1]; p += kBlock) out[p] = M.val[p] * dr * dcol[M.idx[p]];
  } else if (nr > 1 && nr <= kBlock) {
    const int32_t e0 = M.ptr[r0];
    const int32_t cnt = M.ptr[r1] - e0;
    if (tid < nr) s_ptr[tid] = M.ptr[r0 + tid] - e0;
    if (tid == 0) s_ptr[nr] = cnt;
    __syncthreads();
    for (int32_t q = tid; q < cnt; q += kBlock) {
      int32_t lo = 0, hi = nr - 1;  // last k with s_ptr[k] <= q
      while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (s_ptr[mid] <= q) lo = mid;
        else hi = mid - 1;
      }
      const int32_t p = e0 + q;
      out[p] = M.val[p] * drow[r0 + lo] * dcol[M.idx[p]];
    }
  } else if (nr > kBlock) {
    for (int32_t r = r0 + tid; r < r1; r += kBlock)
      for (int32_t p = M.ptr[r]; p < M.ptr[r + 1]; ++p) out[p] = M.val[p] * drow[r] * dcol[M.idx[p]];
  }
}

// out = a * x + b * y (any of the inputs may alias the output).
__global__ void k_pdhg_axpby(double a, const double* x, double b, const double* y, double* out, int32_t n) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a * x[i] + (y ? b * y[i] : 0.0);
}

__global__ void k_pdhg_mul(const double* __restrict__ x, const double* __restrict__ d, double* __restrict__ out,
                           int32_t n) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = x[i] * d[i];
}

// Scaled vectors: multiply (mul) or divide (!mul) by d; infinities stay.
__global__ void k_pdhg_scale_vec(const double* __restrict__ v, const double* __restrict__ d, double* __restrict__ out,
                                 int32_t n, int32_t mul) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = mul ? v[i] * d[i] : v[i] / d[i];
}

// v *= 1 / sqrt(sumsq[0]) when sumsq[0] > 0 (the power iteration's
// normalisation, read from device memory so the host need not wait).
__global__ void k_pdhg_normalize(double* __restrict__ v, const double* __restrict__ sumsq, int32_t n) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const double s = sumsq[0];
  if (i < n && s > 0.0) v[i] *= 1.0 / sqrt(s);
}

// Sum of (a - b)^2 and of a^2 (b may be null) into partials [0], [1].
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
  d2 = block_combine<SumOp>(d2, s_w);
  a2 = block_combine<SumOp>(a2, s_w);
  if (threadIdx.x == 0) {
    double* o = partial + static_cast<int64_t>(blockIdx.x) * kPartials;
    o[0] = d2;
    o[1] = a2;
    o[2] = 0.0;
    o[3] = 0.0;
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
    const double v = block_combine<SumOp>(t, s_w);
    if (threadIdx.x == 0) out[k] = v;
  }
}

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

template <class Op>
hipError_t launch_rows(const gpk_rows& R, const Op& op, double* partial, hipStream_t s) {
  if (R.blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rows<Op>, dim3(static_cast<unsigned>(R.blocks)), dim3(kBlock), 0, s, R.M, R.blk, op,
                     partial);
  return hipGetLastError();
}

hipError_t launch_reduce(const double* partial, int32_t blocks, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_pdhg_reduce, dim3(1), dim3(kBlock), 0, s, partial, blocks, out);
  return hipGetLastError();
}

}  // namespace

extern "C" {

int32_t gpk_pdhg_row_blocks(const int32_t* ptr, int32_t rows, int32_t* blk) {
  int32_t nb = 0, r = 0;
  if (blk) blk[0] = 0;
  while (r < rows) {
    if (ptr[r + 1] - ptr[r] > GPK_PDHG_LONG_ROW) {
      ++r;
    } else {
      const int32_t r_start = r;
      int32_t cnt = 0;
      while (r < rows && r - r_start < kBlock) {
        const int32_t len = ptr[r + 1] - ptr[r];
        if (len > GPK_PDHG_LONG_ROW || cnt + len > kEntryBudget) break;
        cnt += len;
        ++r;
      }
    }
    ++nb;
    if (blk) blk[nb] = r;
  }
  return nb;
}

hipError_t gpk_pdhg_iterations(gpk_rows A, gpk_rows AT, const double* c, const double* l, const double* u,
                               const double* AL, const double* AU, double* x, double* xbar, double* x_acc, double* y,
                               double* y_acc, const double* x0, const double* y0, const double* step,
                               int32_t halpern, int32_t iterations, hipStream_t s) {
  for (int32_t k = 0; k < iterations; ++k) {
    hipError_t e;
    if (halpern) {
      e = launch_rows(AT, OpPrimal<true>{y, c, l, u, step, x, xbar, x_acc, x0, k}, nullptr, s);
      if (e == hipSuccess) e = launch_rows(A, OpDual<true>{xbar, AL, AU, step, y, y_acc, y0, k}, nullptr, s);
    } else {
      e = launch_rows(AT, OpPrimal<false>{y, c, l, u, step, x, xbar, x_acc, x0, k}, nullptr, s);
      if (e == hipSuccess) e = launch_rows(A, OpDual<false>{xbar, AL, AU, step, y, y_acc, y0, k}, nullptr, s);
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
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

hipError_t gpk_pdhg_spmv(gpk_rows M, const double* x, double* y, hipStream_t s) {
  return launch_rows(M, OpSpmv{x, y}, nullptr, s);
}

hipError_t gpk_pdhg_ruiz(gpk_rows M, double* d, hipStream_t s) { return launch_rows(M, OpRuiz{d}, nullptr, s); }

hipError_t gpk_pdhg_scale_csr(gpk_rows M, const double* drow, const double* dcol, double* out, hipStream_t s) {
  if (M.blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rows_scale, dim3(static_cast<unsigned>(M.blocks)), dim3(kBlock), 0, s, M.M, M.blk, drow, dcol,
                     out);
  return hipGetLastError();
}

hipError_t gpk_pdhg_scale_vec(const double* v, const double* d, double* out, int32_t n, int32_t mul, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pdhg_scale_vec, dim3(grid_for(n)), dim3(kBlock), 0, s, v, d, out, n, mul);
  return hipGetLastError();
}

hipError_t gpk_pdhg_normalize(double* v, const double* sumsq, int32_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pdhg_normalize, dim3(grid_for(n)), dim3(kBlock), 0, s, v, sumsq, n);
  return hipGetLastError();
}

int32_t gpk_pdhg_partials(int32_t blocks) { return (blocks > 0 ? blocks : 1) * kPartials; }

hipError_t gpk_pdhg_row_terms(gpk_rows A, const double* AL, const double* AU, const double* x, const double* y,
                              double* partial, double* out, hipStream_t s) {
  const hipError_t e = launch_rows(A, OpRowTerms{x, AL, AU, y}, partial, s);
  if (e != hipSuccess) return e;
  return launch_reduce(partial, A.blocks > 0 ? A.blocks : 0, out, s);
}

hipError_t gpk_pdhg_col_terms(gpk_rows AT, const double* c, const double* l, const double* u, const double* x,
                              const double* y, double* z, double* partial, double* out, hipStream_t s) {
  const hipError_t e = launch_rows(AT, OpColTerms{y, c, l, u, x, z}, partial, s);
  if (e != hipSuccess) return e;
  return launch_reduce(partial, AT.blocks > 0 ? AT.blocks : 0, out, s);
}

hipError_t gpk_pdhg_sq_terms(const double* a, const double* b, int32_t n, double* partial, double* out,
                             hipStream_t s) {
  const unsigned blocks = grid_for(n);
  hipLaunchKernelGGL(k_pdhg_sq_terms, dim3(blocks), dim3(kBlock), 0, s, a, b, n, partial);
  return launch_reduce(partial, static_cast<int32_t>(blocks), out, s);
}

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
