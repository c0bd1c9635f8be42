// Seeded LP generator on the device (no reference counterpart): the request
// stream of the serving workload and the LPs of the batched workload are
// produced where they are consumed, at HBM speed, instead of being built by
// numpy and uploaded.  Every draw is a counter-based hash of (seed, field,
// index), so an LP depends only on its spec: the same spec gives the same
// LP bit for bit, in any process, alone or next to other work.
//
// Distributions (those of the planted synthetic LPs): coefficients +-U[0.5,2]
// rounded to 1/8; c ~ N(0,1), u ~ U[1,10], x0 and the row-bound offsets
// U[0,1] rounded to 2^-10 (every product a*x0 and every row sum is exact);
// l = 0; u = +inf on 20% of the columns, only where c >= 0.  Column classes:
// 1% empty, 2% fixed (l = u = x0), 5% singleton columns (one entry in a
// random general row), the rest normal.  Row classes: 5% singleton rows, 3%
// doubleton equalities, 2% duplicates of a general row scaled by +-2^k
// (k in -2..2), the rest general (length U[len_lo, len_hi], or
// max(min_len, Poisson(mean_len))).  Row kinds: 30% equalities, 10%
// one-sided, the rest ranged; general entries pick normal or fixed columns
// uniformly; a column repeated within a row keeps its first draw.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "gp_device.hpp"
#include "gp_kernels.h"

using namespace gpdev;

namespace {

enum : uint32_t {
  kColClass = 1, kColC1, kColC2, kColU, kColInf, kColX0,
  kRowClass, kRowKind, kRowLen, kRowSide, kRowLo, kRowHi, kDupSrc, kDupFac, kScolRow, kScolVal,
  kEntCol, kEntVal, kSrowCol, kSrowVal, kDblCol, kDblVal
};
enum : uint8_t { kColNormal = 0, kColEmpty = 1, kColFixed = 2, kColSingleton = 3 };
enum : uint8_t { kRowGeneral = 0, kRowSingleton = 1, kRowDoubleton = 2, kRowDuplicate = 3 };
enum : uint8_t { kRanged = 0, kEquality = 1, kOneSided = 2 };
constexpr int kProbeMax = 4096;  // linear probe for a column or row of a wanted class

__device__ inline uint64_t mix64(uint64_t z) {
  // splitmix64 finalizer (Steele, Lea, Flood, OOPSLA 2014)
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ inline uint64_t draw(uint64_t seed, uint32_t field, uint64_t i) {
  return mix64(mix64(seed ^ (static_cast<uint64_t>(field) << 56)) ^ i);
}
// Uniform [0, 1) with 53 random bits.
__device__ inline double unif(uint64_t h) { return static_cast<double>(h >> 11) * 0x1.0p-53; }
__device__ inline double dyadic(double x) { return rint(x * 1024.0) * (1.0 / 1024.0); }
__device__ inline double coeff(uint64_t h) {
  const double mag = rint((0.5 + 1.5 * unif(h)) * 8.0) * 0.125;
  return (h & 1) ? -mag : mag;
}
__device__ inline uint64_t entry_key(int32_t i, int32_t k) {
  return (static_cast<uint64_t>(static_cast<uint32_t>(i)) << 32) | static_cast<uint32_t>(k);
}

__global__ void k_gen_cols(gpk_lpgen_spec sp, uint8_t* cls, double* c, double* l, double* u, double* x0) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= sp.n) return;
  const double uc = unif(draw(sp.seed, kColClass, j));
  const uint8_t k = uc < 0.01 ? kColEmpty : uc < 0.03 ? kColFixed : uc < 0.08 ? kColSingleton : kColNormal;
  // Box-Muller normal, rounded to the 2^-10 grid.
  const double u1 = 1.0 - unif(draw(sp.seed, kColC1, j));
  const double u2 = unif(draw(sp.seed, kColC2, j));
  const double cj = dyadic(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
  double uj = dyadic(1.0 + 9.0 * unif(draw(sp.seed, kColU, j)));
  if (cj >= 0.0 && unif(draw(sp.seed, kColInf, j)) < 0.4) uj = INFINITY;
  const double xj = dyadic(unif(draw(sp.seed, kColX0, j)) * (isfinite(uj) ? uj : 10.0));
  cls[j] = k;
  c[j] = cj;
  x0[j] = xj;
  if (k == kColFixed) {
    l[j] = xj;
    u[j] = xj;
  } else {
    l[j] = 0.0;
    u[j] = uj;
  }
}

// A normal or fixed column (general, singleton-row and doubleton entries).
__device__ inline int32_t pick_column(const gpk_lpgen_spec& sp, const uint8_t* cls, uint64_t h) {
  int32_t j = static_cast<int32_t>(h % static_cast<uint64_t>(sp.n));
  for (int t = 0; t < kProbeMax && (cls[j] == kColEmpty || cls[j] == kColSingleton); ++t) j = j + 1 == sp.n ? 0 : j + 1;
  return j;
}

__device__ inline int32_t poisson(double mean, uint64_t h) {
  // inversion of the Poisson CDF with one uniform draw
  const double v = unif(h);
  double p = exp(-mean), cdf = p;
  int32_t k = 0;
  while (v >= cdf && k < 1000) {
    ++k;
    p *= mean / k;
    cdf += p;
  }
  return k;
}

// Row classes, kinds and raw entry counts (general draws, singleton and
// doubleton entries; singleton-column entries are added by k_gen_scol_count).
__global__ void k_gen_rows(gpk_lpgen_spec sp, uint8_t* rcls, uint8_t* kind, int32_t* raw_count) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sp.m) return;
  const double ur = unif(draw(sp.seed, kRowClass, i));
  const uint8_t rc = ur < 0.05 ? kRowSingleton : ur < 0.08 ? kRowDoubleton : ur < 0.10 ? kRowDuplicate : kRowGeneral;
  const double uk = unif(draw(sp.seed, kRowKind, i));
  uint8_t kd = uk < 0.30 ? kEquality : uk < 0.40 ? kOneSided : kRanged;
  if (rc == kRowDoubleton) kd = kEquality;
  if (rc == kRowSingleton && kd == kOneSided) kd = kRanged;
  int32_t len = 0;
  if (rc == kRowGeneral) {
    const uint64_t h = draw(sp.seed, kRowLen, i);
    if (sp.mean_len > 0.0) {
      len = poisson(sp.mean_len, h);
      if (len < sp.min_len) len = sp.min_len;
    } else {
      len = sp.len_lo + static_cast<int32_t>(h % static_cast<uint64_t>(sp.len_hi - sp.len_lo + 1));
    }
  } else if (rc == kRowSingleton) {
    len = 1;
  } else if (rc == kRowDoubleton) {
    len = 2;
  }
  rcls[i] = rc;
  kind[i] = kd;
  raw_count[i] = len;
}

// A general row for a singleton column or a duplicate row's source; -1 when
// the probe finds none.
__device__ inline int32_t pick_general_row(const gpk_lpgen_spec& sp, const uint8_t* rcls, uint64_t h) {
  int32_t r = static_cast<int32_t>(h % static_cast<uint64_t>(sp.m));
  for (int t = 0; t < kProbeMax; ++t) {
    if (rcls[r] == kRowGeneral) return r;
    r = r + 1 == sp.m ? 0 : r + 1;
  }
  return -1;
}

// Singleton columns: the row each one lands in (-1: none) and its count.
__global__ void k_gen_scol_count(gpk_lpgen_spec sp, const uint8_t* cls, const uint8_t* rcls, int32_t* scol_row,
                                 int32_t* extra) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= sp.n) return;
  int32_t r = -1;
  if (cls[j] == kColSingleton) {
    r = pick_general_row(sp, rcls, draw(sp.seed, kScolRow, j));
    if (r >= 0) atomicAdd(extra + r, 1);
  }
  scol_row[j] = r;
}

__global__ void k_gen_add(int32_t n, int32_t* a, const int32_t* b) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] += b[i];
}

// Raw entries of each row at raw_ptr[i]: its own draws first (in draw
// order); the singleton-column entries go behind them through per-row
// cursors (their order is fixed later by the column sort: their columns are
// distinct and never drawn by a general row).
__global__ void k_gen_fill(gpk_lpgen_spec sp, const uint8_t* cls, const uint8_t* rcls, const int32_t* raw_ptr,
                           int32_t* raw_col, double* raw_val, int32_t* own) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sp.m) return;
  const int32_t b = raw_ptr[i];
  int32_t w = b;
  const uint8_t rc = rcls[i];
  if (rc == kRowGeneral) {
    const uint64_t h = draw(sp.seed, kRowLen, i);
    int32_t len;
    if (sp.mean_len > 0.0) {
      len = poisson(sp.mean_len, h);
      if (len < sp.min_len) len = sp.min_len;
    } else {
      len = sp.len_lo + static_cast<int32_t>(h % static_cast<uint64_t>(sp.len_hi - sp.len_lo + 1));
    }
    for (int32_t k = 0; k < len; ++k) {
      raw_col[w] = pick_column(sp, cls, draw(sp.seed, kEntCol, entry_key(i, k)));
      raw_val[w] = coeff(draw(sp.seed, kEntVal, entry_key(i, k)));
      ++w;
    }
  } else if (rc == kRowSingleton) {
    raw_col[w] = pick_column(sp, cls, draw(sp.seed, kSrowCol, i));
    raw_val[w] = coeff(draw(sp.seed, kSrowVal, i));
    ++w;
  } else if (rc == kRowDoubleton) {
    const int32_t c1 = pick_column(sp, cls, draw(sp.seed, kDblCol, entry_key(i, 0)));
    int32_t c2 = pick_column(sp, cls, draw(sp.seed, kDblCol, entry_key(i, 1)));
    if (c2 == c1) c2 = pick_column(sp, cls, static_cast<uint64_t>(c1 + 1 == sp.n ? 0 : c1 + 1));
    raw_col[w] = c1;
    raw_val[w] = coeff(draw(sp.seed, kDblVal, entry_key(i, 0)));
    raw_col[w + 1] = c2;
    raw_val[w + 1] = coeff(draw(sp.seed, kDblVal, entry_key(i, 1)));
    w += 2;
  }
  own[i] = w - b;
}

__global__ void k_gen_scol_fill(gpk_lpgen_spec sp, const int32_t* scol_row, const int32_t* raw_ptr, const int32_t* own,
                                int32_t* cursor, int32_t* raw_col, double* raw_val) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= sp.n) return;
  const int32_t r = scol_row[j];
  if (r < 0) return;
  const int32_t p = raw_ptr[r] + own[r] + atomicAdd(cursor + r, 1);
  raw_col[p] = j;
  raw_val[p] = coeff(draw(sp.seed, kScolVal, j));
}

// Per row: stable insertion sort of the raw segment by column, first draw of
// a repeated column kept; unique[i] = distinct entries.  Duplicate rows pick
// their source general row (src[i], -1 for other rows).
__global__ void k_gen_sort_rows(gpk_lpgen_spec sp, const uint8_t* rcls, const int32_t* raw_ptr, int32_t* raw_col,
                                double* raw_val, int32_t* unique, int32_t* src) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sp.m) return;
  const int32_t b = raw_ptr[i], e = raw_ptr[i + 1];
  for (int32_t p = b + 1; p < e; ++p) {
    const int32_t kc = raw_col[p];
    const double kv = raw_val[p];
    int32_t q = p - 1;
    while (q >= b && raw_col[q] > kc) {
      raw_col[q + 1] = raw_col[q];
      raw_val[q + 1] = raw_val[q];
      --q;
    }
    raw_col[q + 1] = kc;
    raw_val[q + 1] = kv;
  }
  int32_t u = 0;
  for (int32_t p = b; p < e; ++p)
    if (p == b || raw_col[p] != raw_col[p - 1]) ++u;
  unique[i] = u;
  src[i] = rcls[i] == kRowDuplicate ? pick_general_row(sp, rcls, draw(sp.seed, kDupSrc, i)) : -1;
}

__device__ inline double dup_factor(const gpk_lpgen_spec& sp, int32_t i) {
  const uint64_t h = draw(sp.seed, kDupFac, i);
  const int k = static_cast<int>(h % 5) - 2;
  const double f = ldexp(1.0, k);
  return ((h >> 8) & 1) ? -f : f;
}

__global__ void k_gen_lengths(int32_t m, const int32_t* unique, const int32_t* src, int32_t* len) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  len[i] = src[i] >= 0 ? unique[src[i]] : (src[i] == -1 ? unique[i] : 0);
}

// Final CSR: a row's distinct entries, or its source row's scaled by +-2^k.
__global__ void k_gen_write(gpk_lpgen_spec sp, const uint8_t* rcls, const int32_t* raw_ptr, const int32_t* raw_col,
                            const double* raw_val, const int32_t* src, const int32_t* ptr, int32_t* idx, double* val) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sp.m) return;
  const bool dup = rcls[i] == kRowDuplicate;
  if (dup && src[i] < 0) return;  // no general row found: an empty row
  const int32_t s = dup ? src[i] : i;
  const double f = dup ? dup_factor(sp, i) : 1.0;
  int32_t w = ptr[i];
  const int32_t b = raw_ptr[s], e = raw_ptr[s + 1];
  for (int32_t p = b; p < e; ++p) {
    if (p > b && raw_col[p] == raw_col[p - 1]) continue;
    idx[w] = raw_col[p];
    val[w] = raw_val[p] * f;
    ++w;
  }
}

// Row bounds from the exact activity a . x0 (duplicate rows in a second
// pass, from their source's bounds).
__global__ void k_gen_bounds(gpk_lpgen_spec sp, int32_t pass, const uint8_t* rcls, const uint8_t* kind,
                             const int32_t* src, const int32_t* ptr, const int32_t* idx, const double* val,
                             const double* x0, double* AL, double* AU) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sp.m) return;
  const bool dup = rcls[i] == kRowDuplicate && src[i] >= 0;
  if ((pass == 0) == dup) return;
  if (dup) {
    const double f = dup_factor(sp, i);
    const double a = AL[src[i]] * f, b = AU[src[i]] * f;
    AL[i] = f > 0.0 ? a : b;
    AU[i] = f > 0.0 ? b : a;
    return;
  }
  double ax = 0.0;
  for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) ax += val[p] * x0[idx[p]];
  const double lo = ax - dyadic(unif(draw(sp.seed, kRowLo, i)));
  const double hi = ax + dyadic(unif(draw(sp.seed, kRowHi, i)));
  const uint8_t kd = kind[i];
  if (kd == kEquality) {
    AL[i] = ax;
    AU[i] = ax;
  } else if (kd == kOneSided) {
    const bool side = unif(draw(sp.seed, kRowSide, i)) < 0.5;
    AL[i] = side ? -INFINITY : lo;
    AU[i] = side ? hi : INFINITY;
  } else {
    AL[i] = lo;
    AU[i] = hi;
  }
}

}  // namespace

#define GPK_GEN_LAUNCH(kernel, count, ...)                                                       \
  do {                                                                                           \
    if ((count) > 0) hipLaunchKernelGGL(kernel, dim3(grid_for(count)), dim3(kBlock), 0, s, __VA_ARGS__); \
    e = hipGetLastError();                                                                       \
    if (e != hipSuccess) return e;                                                               \
  } while (0)

extern "C" {

hipError_t gpk_lpgen_columns(gpk_lpgen_spec sp, uint8_t* col_class, double* c, double* l, double* u, double* x0,
                             hipStream_t s) {
  hipError_t e = hipSuccess;
  GPK_GEN_LAUNCH(k_gen_cols, sp.n, sp, col_class, c, l, u, x0);
  return e;
}

hipError_t gpk_lpgen_row_counts(gpk_lpgen_spec sp, const uint8_t* col_class, uint8_t* row_class, uint8_t* kind,
                                int32_t* raw_count, int32_t* scol_row, int32_t* extra, hipStream_t s) {
  hipError_t e = hipSuccess;
  GPK_GEN_LAUNCH(k_gen_rows, sp.m, sp, row_class, kind, raw_count);
  GPK_GEN_LAUNCH(k_gen_scol_count, sp.n, sp, col_class, row_class, scol_row, extra);
  GPK_GEN_LAUNCH(k_gen_add, sp.m, sp.m, raw_count, static_cast<const int32_t*>(extra));
  return e;
}

hipError_t gpk_lpgen_fill(gpk_lpgen_spec sp, const uint8_t* col_class, const uint8_t* row_class,
                          const int32_t* raw_ptr, const int32_t* scol_row, int32_t* own, int32_t* cursor,
                          int32_t* raw_col, double* raw_val, int32_t* unique, int32_t* src, int32_t* len,
                          hipStream_t s) {
  hipError_t e = hipSuccess;
  GPK_GEN_LAUNCH(k_gen_fill, sp.m, sp, col_class, row_class, raw_ptr, raw_col, raw_val, own);
  GPK_GEN_LAUNCH(k_gen_scol_fill, sp.n, sp, scol_row, raw_ptr, static_cast<const int32_t*>(own), cursor, raw_col,
                 raw_val);
  GPK_GEN_LAUNCH(k_gen_sort_rows, sp.m, sp, row_class, raw_ptr, raw_col, raw_val, unique, src);
  GPK_GEN_LAUNCH(k_gen_lengths, sp.m, sp.m, static_cast<const int32_t*>(unique), static_cast<const int32_t*>(src),
                 len);
  return e;
}

hipError_t gpk_lpgen_write(gpk_lpgen_spec sp, const uint8_t* row_class, const uint8_t* kind, const int32_t* raw_ptr,
                           const int32_t* raw_col, const double* raw_val, const int32_t* src, const int32_t* ptr,
                           int32_t* idx, double* val, const double* x0, double* AL, double* AU, hipStream_t s) {
  hipError_t e = hipSuccess;
  GPK_GEN_LAUNCH(k_gen_write, sp.m, sp, row_class, raw_ptr, raw_col, raw_val, src, ptr, idx, val);
  GPK_GEN_LAUNCH(k_gen_bounds, sp.m, sp, 0, row_class, kind, src, ptr, static_cast<const int32_t*>(idx),
                 static_cast<const double*>(val), x0, AL, AU);
  GPK_GEN_LAUNCH(k_gen_bounds, sp.m, sp, 1, row_class, kind, src, ptr, static_cast<const int32_t*>(idx),
                 static_cast<const double*>(val), x0, AL, AU);
  return e;
}

}  // extern "C"
