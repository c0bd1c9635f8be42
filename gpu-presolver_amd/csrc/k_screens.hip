// Fail-closed screens of the six structure-specific rules (covering-cost
// dominance, Linf components, antipodal components, bounded two-row
// projection, orphan McCormick projection, structural L1 substitution).
// This library does not rewrite those block structures; instead each screen
// evaluates the test that decides whether the reference's rule would act on
// the LP, and the scheduler stops with an error when it would, so a record
// never silently differs from the reference's.  Every decision below is a
// comparison or a short sum over a handful of entries, one thread per row,
// column or sampled pair.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "gp_device.hpp"
#include "gp_kernels.h"

using namespace gpdev;

namespace {

__device__ inline bool pos_inf(double v) { return isinf(v) && v > 0.0; }
__device__ inline bool neg_inf(double v) { return isinf(v) && v < 0.0; }

// ------------------------------------------------ covering-cost dominance ----
// rule_covering_cost_dominance.cu:112-196 (full validation) and 173-182
// (dominated columns): every row 1 <= a x < +inf with unit entries and a
// cost-1 column, every column binary with a positive integer cost; column j
// is dominated when c_j exceeds its row count.
__global__ void k_scr_covering_rows(const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                    const double* __restrict__ c, const double* __restrict__ AL,
                                    const double* __restrict__ AU, int32_t m, int32_t* fail) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  bool ok = AL[i] == 1.0 && pos_inf(AU[i]);
  bool witness = false;
  for (int32_t p = ptr[i]; p < ptr[i + 1] && !witness; ++p) witness = c[idx[p]] == 1.0;
  if (!(ok && witness)) wave_set_flag(fail);
}

__global__ void k_scr_covering_cols(const double* __restrict__ c, const double* __restrict__ l,
                                    const double* __restrict__ u, const int32_t* __restrict__ at_ptr, int32_t n,
                                    int32_t* fail, int32_t* dominated) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double cj = c[j];
  const bool ok = l[j] == 0.0 && u[j] == 1.0 && isfinite(cj) && cj >= 1.0 && floor(cj) == cj;
  if (!ok) wave_set_flag(fail);
  if (cj > static_cast<double>(at_ptr[j + 1] - at_ptr[j])) wave_add1(dominated);
}

__global__ void k_scr_unit_values(const double* __restrict__ v, int32_t nnz, int32_t* fail) {
  const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nnz && v[p] != 1.0) wave_set_flag(fail);
}

// ------------------------------------------------ bounded two-row projection --
// rule_bounded_two_row_projection.cu:18-296: a zero-cost column with a
// finite box, in exactly two rows, each row one-sided, at most 8 entries and
// giving the column a bound of opposite sense as an affine function of the
// row's other columns; the projection is safe when, over the box of those
// columns, the lower affine stays below u_j, the upper stays above l_j and
// the lower minus the upper never exceeds 0, each with the relative guard
// tol.  These sums are formed without fused multiply-adds, so a build of the
// rule that fuses them may differ in the last bit; every guard is therefore
// tested twice, at tol / 2 (bit 1: the rule may take the column) and at 2 tol
// (bit 2: it surely takes it), and the count built from the two bits bounds
// the rule's own count from above.
constexpr int kBtrMaxRow = 8;
constexpr int kBtrTerms = kBtrMaxRow - 1;
constexpr int kBtrDiffTerms = 2 * kBtrTerms;

struct Affine {
  double constant;
  int32_t len;
  int32_t col[kBtrTerms];
  double coef[kBtrTerms];
};

// Adds coef * x_col to a term list, merging a repeated column.
__device__ inline bool add_term(int32_t col, double coef, int32_t* cols, double* coefs, int32_t* len, int32_t cap) {
  if (!isfinite(coef)) return false;
  for (int32_t k = 0; k < *len; ++k) {
    if (cols[k] == col) {
      coefs[k] += coef;
      return isfinite(coefs[k]);
    }
  }
  if (*len >= cap) return false;
  cols[*len] = col;
  coefs[*len] = coef;
  ++*len;
  return true;
}

// lhs <= rhs with margin tol * (1 + max |.|): bit 1 at tol / 2, bit 2 at 2 tol.
__device__ inline uint8_t guarded_le(double lhs, double rhs, double tol) {
  if (!isfinite(lhs) || !isfinite(rhs)) return 0;
  const double d = rhs - lhs, g = 1.0 + fmax(fabs(lhs), fabs(rhs));
  return static_cast<uint8_t>((d >= 0.5 * tol * g ? 1 : 0) | (d >= 2.0 * tol * g ? 2 : 0));
}

// Row `row` solved for column j: x_j >= (lower) or <= (upper) constant + sum coef x_col.
__device__ bool row_affine(Affine* a, bool* lower, const int32_t* ptr, const int32_t* idx, const double* val,
                           const double* AL, const double* AU, int32_t row, int32_t j, int32_t max_row, double ztol) {
  const int32_t b = ptr[row], e = ptr[row + 1];
  if (e - b < 2 || e - b > max_row) return false;
  int32_t hits = 0;
  double piv = 0.0;
  for (int32_t p = b; p < e; ++p)
    if (idx[p] == j) {
      ++hits;
      piv = val[p];
    }
  if (hits != 1 || !isfinite(piv) || fabs(piv) <= ztol) return false;
  double bound;
  if (isfinite(AL[row]) && pos_inf(AU[row])) {
    *lower = piv > 0.0;
    bound = AL[row];
  } else if (neg_inf(AL[row]) && isfinite(AU[row])) {
    *lower = piv < 0.0;
    bound = AU[row];
  } else {
    return false;
  }
  a->constant = bound / piv;
  a->len = 0;
  if (!isfinite(a->constant)) return false;
  for (int32_t p = b; p < e; ++p)
    if (idx[p] != j && !add_term(idx[p], -val[p] / piv, a->col, a->coef, &a->len, kBtrTerms)) return false;
  return true;
}

__device__ bool affine_range(const Affine& a, const double* l, const double* u, double* lo, double* hi) {
  double mn = a.constant, mx = a.constant;
  for (int32_t k = 0; k < a.len; ++k) {
    const double lk = l[a.col[k]], uk = u[a.col[k]], q = a.coef[k];
    if (!isfinite(lk) || !isfinite(uk) || lk > uk) return false;
    mn += q >= 0.0 ? q * lk : q * uk;
    mx += q >= 0.0 ? q * uk : q * lk;
  }
  if (!isfinite(mn) || !isfinite(mx)) return false;
  *lo = mn;
  *hi = mx;
  return true;
}

__global__ void k_scr_btr_candidates(uint8_t* __restrict__ cand, const int32_t* __restrict__ ptr,
                                     const int32_t* __restrict__ idx, const double* __restrict__ val,
                                     const int32_t* __restrict__ at_ptr, const int32_t* __restrict__ at_idx,
                                     const double* __restrict__ c, const double* __restrict__ AL,
                                     const double* __restrict__ AU, const double* __restrict__ l,
                                     const double* __restrict__ u, int32_t n, int32_t max_row, double ztol,
                                     double tol) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  cand[j] = 0;
  if (c[j] != 0.0 || !isfinite(l[j]) || !isfinite(u[j]) || at_ptr[j + 1] - at_ptr[j] != 2) return;
  uint8_t bits = guarded_le(l[j], u[j], tol);
  if (bits == 0) return;
  const int32_t r0 = at_idx[at_ptr[j]], r1 = at_idx[at_ptr[j] + 1];
  if (r0 == r1) return;
  Affine a0, a1;
  bool lo0 = false, lo1 = false;
  if (!row_affine(&a0, &lo0, ptr, idx, val, AL, AU, r0, j, max_row, ztol) ||
      !row_affine(&a1, &lo1, ptr, idx, val, AL, AU, r1, j, max_row, ztol) || lo0 == lo1)
    return;
  const Affine& lower = lo0 ? a0 : a1;
  const Affine& upper = lo0 ? a1 : a0;
  double lower_min, lower_max, upper_min, upper_max;
  if (!affine_range(lower, l, u, &lower_min, &lower_max) || !affine_range(upper, l, u, &upper_min, &upper_max))
    return;
  bits &= guarded_le(lower_max, u[j], tol) & guarded_le(l[j], upper_min, tol);
  if (bits == 0) return;
  int32_t cols[kBtrDiffTerms];
  double coefs[kBtrDiffTerms];
  int32_t len = 0;
  for (int32_t k = 0; k < lower.len; ++k)
    if (!add_term(lower.col[k], lower.coef[k], cols, coefs, &len, kBtrDiffTerms)) return;
  for (int32_t k = 0; k < upper.len; ++k)
    if (!add_term(upper.col[k], -upper.coef[k], cols, coefs, &len, kBtrDiffTerms)) return;
  double diff = lower.constant - upper.constant;
  for (int32_t k = 0; k < len; ++k) {
    const double lk = l[cols[k]], uk = u[cols[k]];
    if (!isfinite(lk) || !isfinite(uk) || lk > uk) return;
    diff += coefs[k] >= 0.0 ? coefs[k] * uk : coefs[k] * lk;
  }
  if (!isfinite(diff)) return;
  bits &= guarded_le(diff, 0.0, tol);
  cand[j] = bits & 2 ? 3 : bits;  // a sure candidate is also a possible one
}

// Rows touched by sure candidates (cand bit 2; deg entries per column).
__global__ void k_scr_row_hits(const uint8_t* __restrict__ cand, const int32_t* __restrict__ at_ptr,
                               const int32_t* __restrict__ at_idx, int32_t n, int32_t deg, int32_t* row_hits) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || (cand[j] & 2) == 0) return;
  for (int32_t k = 0; k < deg; ++k) atomicAdd(&row_hits[at_idx[at_ptr[j] + k]], 1);
}

// The rule keeps a candidate whose rows no other candidate touches (and, with
// endpoints, whose endpoint columns are not candidates).  Counted here: every
// possible candidate (bit 1) that can still be alone on its rows once it is
// counted itself (a sure one is already in row_hits), so the count is at
// least the rule's.
__global__ void k_scr_select(const uint8_t* __restrict__ cand, const int32_t* __restrict__ at_ptr,
                             const int32_t* __restrict__ at_idx, const int32_t* __restrict__ row_hits, int32_t n,
                             int32_t deg, const int32_t* __restrict__ end0, const int32_t* __restrict__ end1,
                             int32_t* selected) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || (cand[j] & 1) == 0) return;
  const int32_t own = cand[j] & 2 ? 0 : 1;
  bool ok = true;
  for (int32_t k = 0; k < deg; ++k) ok = ok && row_hits[at_idx[at_ptr[j] + k]] + own <= 1;
  if (end0 != nullptr) ok = ok && (cand[end0[j]] & 1) == 0 && (cand[end1[j]] & 1) == 0;
  if (ok) wave_add1(selected);
}

// ------------------------------------------------ orphan McCormick projection --
// rule_orphan_mccormick_projection.cu:50-195: z binary with zero cost in
// exactly three rows: z >= x + y - 1 (three entries, unit magnitudes equal to
// the pivot's) and z <= x, z <= y (two entries each), x != y binary.
struct McRow {
  bool ok;
  bool lower;
  int32_t e0, e1;
};

__device__ McRow mc_row(const int32_t* ptr, const int32_t* idx, const double* val, const double* AL,
                        const double* AU, int32_t row, int32_t z) {
  McRow r{false, false, -1, -1};
  const int32_t b = ptr[row], e = ptr[row + 1];
  if (e - b != 2 && e - b != 3) return r;
  double piv = 0.0;
  int32_t hits = 0;
  for (int32_t p = b; p < e; ++p)
    if (idx[p] == z) {
      piv = val[p];
      ++hits;
    }
  if (hits != 1 || !isfinite(piv) || piv == 0.0) return r;
  double bound;
  if (isfinite(AL[row]) && pos_inf(AU[row])) {
    r.lower = piv > 0.0;
    bound = AL[row];
  } else if (neg_inf(AL[row]) && isfinite(AU[row])) {
    r.lower = piv < 0.0;
    bound = AU[row];
  } else {
    return r;
  }
  if (e - b != (r.lower ? 3 : 2) || bound != (r.lower ? -piv : 0.0)) return r;
  int32_t ends = 0;
  for (int32_t p = b; p < e; ++p) {
    if (idx[p] == z) continue;
    if (val[p] != -piv || ends >= 2) return r;
    if (ends == 0) r.e0 = idx[p];
    else r.e1 = idx[p];
    ++ends;
  }
  if (ends != (r.lower ? 2 : 1) || (r.lower && r.e0 == r.e1)) return r;
  r.ok = true;
  return r;
}

__global__ void k_scr_mccormick_candidates(uint8_t* __restrict__ cand, int32_t* __restrict__ end0,
                                           int32_t* __restrict__ end1, const int32_t* __restrict__ ptr,
                                           const int32_t* __restrict__ idx, const double* __restrict__ val,
                                           const int32_t* __restrict__ at_ptr, const int32_t* __restrict__ at_idx,
                                           const double* __restrict__ c, const double* __restrict__ AL,
                                           const double* __restrict__ AU, const double* __restrict__ l,
                                           const double* __restrict__ u, int32_t n) {
  const int32_t z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= n) return;
  cand[z] = 0;
  end0[z] = -1;
  end1[z] = -1;
  if (c[z] != 0.0 || l[z] != 0.0 || u[z] != 1.0 || at_ptr[z + 1] - at_ptr[z] != 3) return;
  const int32_t* rows = at_idx + at_ptr[z];
  if (rows[0] == rows[1] || rows[0] == rows[2] || rows[1] == rows[2]) return;
  int32_t lower = -1, up[2] = {-1, -1}, n_up = 0;
  McRow pat[3];
  for (int k = 0; k < 3; ++k) {
    pat[k] = mc_row(ptr, idx, val, AL, AU, rows[k], z);
    if (!pat[k].ok) return;
    if (pat[k].lower) {
      if (lower >= 0) return;
      lower = k;
    } else {
      if (n_up >= 2) return;
      up[n_up++] = k;
    }
  }
  if (lower < 0 || n_up != 2) return;
  const int32_t x = pat[lower].e0, y = pat[lower].e1;
  const int32_t ux = pat[up[0]].e0, uy = pat[up[1]].e0;
  if (x < 0 || y < 0 || x == y || !((ux == x && uy == y) || (ux == y && uy == x)) || l[x] != 0.0 || u[x] != 1.0 ||
      l[y] != 0.0 || u[y] != 1.0)
    return;
  end0[z] = x;
  end1[z] = y;
  cand[z] = 3;  // exact comparisons only: possible and sure alike
}

// ------------------------------------------------ antipodal / Linf quick probes --
// rule_antipodal_components.cu:41-125 and rule_linf_components.cu:108-220:
// sampled column pairs (j, j + pairs) must mirror each other, and at least 5%
// of sampled rows must be merge rows a x_p - a x_q - a x_{p+P} + a x_{q+P} = 0.
__device__ inline int32_t sampled(int32_t s, int32_t samples, int32_t count) {
  return static_cast<int32_t>((static_cast<int64_t>(s) * count) / samples);
}

__global__ void k_scr_antipodal_pairs(const int32_t* __restrict__ at_ptr, const int32_t* __restrict__ at_idx,
                                      const double* __restrict__ at_val, const double* __restrict__ c,
                                      const double* __restrict__ l, const double* __restrict__ u, int32_t pairs,
                                      int32_t probes, int32_t* fail) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= probes) return;
  const int32_t j = sampled(s, probes, pairs), q = j + pairs;
  const int32_t b = at_ptr[j], d = at_ptr[j + 1] - b, qb = at_ptr[q];
  bool ok = d >= 1 && d <= 64 && d == at_ptr[q + 1] - qb && isfinite(c[j]) && c[j] > 0.0 && c[j] == c[q];
  for (int32_t k : {j, q}) ok = ok && isfinite(l[k]) && !isnan(u[k]) && u[k] >= l[k];
  ok = ok && (!isfinite(u[j]) || isfinite(u[j] - l[j])) && (!isfinite(u[q]) || isfinite(u[q] - l[q]));
  for (int32_t p = 0; ok && p < d; ++p) ok = at_idx[b + p] == at_idx[qb + p] && at_val[b + p] == -at_val[qb + p];
  if (!ok) wave_set_flag(fail);
}

__global__ void k_scr_linf_pairs(const int32_t* __restrict__ at_ptr, const int32_t* __restrict__ at_idx,
                                 const double* __restrict__ at_val, const int32_t* __restrict__ ptr,
                                 const int32_t* __restrict__ idx, const double* __restrict__ val,
                                 const double* __restrict__ AL, const double* __restrict__ AU,
                                 const double* __restrict__ c, const double* __restrict__ l,
                                 const double* __restrict__ u, int32_t pairs, int32_t n, int32_t probes,
                                 int32_t* fail) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= probes) return;
  const int32_t j = sampled(s, probes, pairs), q = j + pairs;
  const int32_t b = at_ptr[j], d = at_ptr[j + 1] - b, qb = at_ptr[q];
  bool ok = d >= 2 && d <= 128 && d == at_ptr[q + 1] - qb && c[j] == 0.0 && c[q] == 0.0;
  for (int32_t k : {j, q})
    ok = ok && isfinite(l[k]) && !isnan(u[k]) && u[k] >= l[k] && (!isfinite(u[k]) || isfinite(u[k] - l[k]));
  int32_t same = 0;
  for (int32_t p = 0; ok && p < d; ++p) {
    const int32_t r = at_idx[b + p];
    const double a = at_val[b + p], aq = at_val[qb + p];
    if (r != at_idx[qb + p] || !isfinite(a) || !isfinite(aq) || a == 0.0) {
      ok = false;
    } else if (a == aq) {
      ++same;
      const int32_t rb = ptr[r];
      ok = ptr[r + 1] - rb == 3 && idx[rb] == j && idx[rb + 1] == q && idx[rb + 2] >= 2 * pairs &&
           idx[rb + 2] < n && val[rb] == val[rb + 1] && val[rb] < 0.0 && isfinite(val[rb + 2]) &&
           val[rb + 2] > 0.0 && isfinite(AL[r]) && AU[r] == INFINITY;
    } else if (a != -aq) {
      ok = false;
    }
  }
  if (!ok || same != 1) wave_set_flag(fail);
}

__global__ void k_scr_merge_rows(const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                 const double* __restrict__ val, const double* __restrict__ AL,
                                 const double* __restrict__ AU, int32_t m, int32_t pairs, int32_t probes,
                                 int32_t* merges) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= probes) return;
  const int32_t r = sampled(s, probes, m);
  const int32_t b = ptr[r];
  if (ptr[r + 1] - b != 4 || AL[r] != 0.0 || AU[r] != 0.0) return;
  const int32_t pa = idx[b], pb = idx[b + 1];
  const double va = val[b], vb = val[b + 1];
  const bool merge = pa >= 0 && pa < pairs && pb >= 0 && pb < pairs && pa != pb && idx[b + 2] == pa + pairs &&
                     idx[b + 3] == pb + pairs && isfinite(va) && isfinite(vb) && va != 0.0 && va == -vb &&
                     va == -val[b + 2] && vb == -val[b + 3];
  if (merge) wave_add1(merges);
}

// ------------------------------------------------ structural L1 prefix screen --
// rule_structural_l1_substitution.cu:2214-2376: the first rows of A must open
// one of the three block patterns (an L1 split, an outer pair, a graph
// block), read from rows 0..3 only.
struct Pre {
  bool eq;
  bool zl;  // two entries, 0 <= a x (or a x <= 0, stored negated)
  int32_t c1, c2;
  double v1, v2;
};

__device__ Pre pre_row(const int32_t* ptr, const int32_t* idx, const double* val, const double* AL, const double* AU,
                       int32_t row) {
  Pre o{false, false, -1, -1, 0.0, 0.0};
  const double al = AL[row], au = AU[row];
  o.eq = isfinite(al) && isfinite(au) && al == au;
  const int32_t b = ptr[row];
  if (ptr[row + 1] - b != 2) return o;
  int32_t c1 = idx[b], c2 = idx[b + 1];
  double v1 = val[b], v2 = val[b + 1];
  if (c2 < c1) {
    const int32_t tc = c1;
    c1 = c2;
    c2 = tc;
    const double tv = v1;
    v1 = v2;
    v2 = tv;
  }
  const bool zero_lower = al == 0.0 && pos_inf(au);
  const bool zero_upper = neg_inf(al) && au == 0.0;
  if (zero_lower || zero_upper) {
    const double sgn = zero_lower ? 1.0 : -1.0;
    o.zl = true;
    o.c1 = c1;
    o.c2 = c2;
    o.v1 = sgn * v1;
    o.v2 = sgn * v2;
  }
  return o;
}

__device__ bool l1_pair(const Pre& a, const Pre& b) {
  if (a.c1 != b.c1 || a.c2 != b.c2) return false;
  const bool first = isfinite(a.v1) && a.v1 > 0.0 && a.v1 == b.v1 && a.v2 == -a.v1 && b.v2 == b.v1;
  const bool second = isfinite(a.v2) && a.v2 > 0.0 && a.v2 == b.v2 && a.v1 == -a.v2 && b.v1 == b.v2;
  return first || second;
}

__global__ void k_scr_structural_prefix(const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                        const double* __restrict__ val, const double* __restrict__ AL,
                                        const double* __restrict__ AU, int32_t m, int32_t* pass) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  bool hit = false;
  if (m >= 3 && m % 3 == 0) {  // L1 split: an equality, then two paired 0 <= rows
    const Pre e = pre_row(ptr, idx, val, AL, AU, 0), r1 = pre_row(ptr, idx, val, AL, AU, 1),
              r2 = pre_row(ptr, idx, val, AL, AU, 2);
    hit = hit || (e.eq && r1.zl && r2.zl && l1_pair(r1, r2));
  }
  if (m >= 11) {  // outer pair: three 0 <= rows over one column pair with signs (+-, ++, -+)
    const Pre a = pre_row(ptr, idx, val, AL, AU, 0), b = pre_row(ptr, idx, val, AL, AU, 1),
              d = pre_row(ptr, idx, val, AL, AU, 2);
    hit = hit || (a.zl && b.zl && d.zl && a.c1 == b.c1 && a.c1 == d.c1 && a.c2 == b.c2 && a.c2 == d.c2 &&
                  a.v1 == 1.0 && a.v2 == -1.0 && b.v1 == 1.0 && b.v2 == 1.0 && d.v1 == -1.0 && d.v2 == 1.0);
  }
  if (m >= 4 && m % 4 == 0) {  // graph block: an equality, a paired couple, then a mixed-sign row sharing a column
    const Pre e = pre_row(ptr, idx, val, AL, AU, 0), r1 = pre_row(ptr, idx, val, AL, AU, 1),
              r2 = pre_row(ptr, idx, val, AL, AU, 2), r3 = pre_row(ptr, idx, val, AL, AU, 3);
    constexpr double kTol = 1.0e-12;
    const bool pos = r3.v1 > kTol || r3.v2 > kTol;
    const bool neg = r3.v1 < -kTol || r3.v2 < -kTol;
    const bool overlap = r3.c1 == r1.c1 || r3.c1 == r1.c2 || r3.c2 == r1.c1 || r3.c2 == r1.c2;
    hit = hit || (e.eq && r1.zl && r2.zl && l1_pair(r1, r2) && r3.zl && pos && neg && overlap);
  }
  *pass = hit ? 1 : 0;
}

}  // namespace

extern "C" {

hipError_t gpk_screen_covering(const int32_t* ptr, const int32_t* idx, const double* val, const int32_t* at_ptr,
                               const double* at_val, const double* c, const double* AL, const double* AU,
                               const double* l, const double* u, int32_t m, int32_t n, int32_t nnz, int32_t* out,
                               hipStream_t s) {
  if (m > 0)
    hipLaunchKernelGGL(k_scr_covering_rows, dim3(grid_for(m)), dim3(kBlock), 0, s, ptr, idx, c, AL, AU, m, out);
  if (n > 0)
    hipLaunchKernelGGL(k_scr_covering_cols, dim3(grid_for(n)), dim3(kBlock), 0, s, c, l, u, at_ptr, n, out,
                       out + 1);
  if (nnz > 0) {
    hipLaunchKernelGGL(k_scr_unit_values, dim3(grid_for(nnz)), dim3(kBlock), 0, s, val, nnz, out);
    hipLaunchKernelGGL(k_scr_unit_values, dim3(grid_for(nnz)), dim3(kBlock), 0, s, at_val, nnz, out);
  }
  return hipGetLastError();
}

hipError_t gpk_screen_bounded_two_row(const int32_t* ptr, const int32_t* idx, const double* val,
                                      const int32_t* at_ptr, const int32_t* at_idx, const double* c,
                                      const double* AL, const double* AU, const double* l, const double* u,
                                      int32_t n, int32_t max_row, double ztol, double tol, uint8_t* cand,
                                      int32_t* row_hits, int32_t* selected, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (max_row > kBtrMaxRow) max_row = kBtrMaxRow;
  hipLaunchKernelGGL(k_scr_btr_candidates, dim3(grid_for(n)), dim3(kBlock), 0, s, cand, ptr, idx, val, at_ptr,
                     at_idx, c, AL, AU, l, u, n, max_row, ztol, tol);
  hipLaunchKernelGGL(k_scr_row_hits, dim3(grid_for(n)), dim3(kBlock), 0, s, cand, at_ptr, at_idx, n, 2, row_hits);
  hipLaunchKernelGGL(k_scr_select, dim3(grid_for(n)), dim3(kBlock), 0, s, cand, at_ptr, at_idx, row_hits, n, 2,
                     nullptr, nullptr, selected);
  return hipGetLastError();
}

hipError_t gpk_screen_mccormick(const int32_t* ptr, const int32_t* idx, const double* val, const int32_t* at_ptr,
                                const int32_t* at_idx, const double* c, const double* AL, const double* AU,
                                const double* l, const double* u, int32_t n, uint8_t* cand, int32_t* ends,
                                int32_t* row_hits, int32_t* selected, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scr_mccormick_candidates, dim3(grid_for(n)), dim3(kBlock), 0, s, cand, ends, ends + n, ptr,
                     idx, val, at_ptr, at_idx, c, AL, AU, l, u, n);
  hipLaunchKernelGGL(k_scr_row_hits, dim3(grid_for(n)), dim3(kBlock), 0, s, cand, at_ptr, at_idx, n, 3, row_hits);
  hipLaunchKernelGGL(k_scr_select, dim3(grid_for(n)), dim3(kBlock), 0, s, cand, at_ptr, at_idx, row_hits, n, 3,
                     ends, ends + n, selected);
  return hipGetLastError();
}

hipError_t gpk_screen_pair_probe(int32_t kind, const int32_t* ptr, const int32_t* idx, const double* val,
                                 const int32_t* at_ptr, const int32_t* at_idx, const double* at_val,
                                 const double* c, const double* AL, const double* AU, const double* l,
                                 const double* u, int32_t m, int32_t n, int32_t pairs, int32_t pair_probes,
                                 int32_t row_probes, int32_t* out, hipStream_t s) {
  if (pair_probes <= 0 || row_probes <= 0) return hipErrorInvalidValue;
  if (kind == 0)
    hipLaunchKernelGGL(k_scr_antipodal_pairs, dim3(grid_for(pair_probes)), dim3(kBlock), 0, s, at_ptr, at_idx,
                       at_val, c, l, u, pairs, pair_probes, out);
  else
    hipLaunchKernelGGL(k_scr_linf_pairs, dim3(grid_for(pair_probes)), dim3(kBlock), 0, s, at_ptr, at_idx, at_val,
                       ptr, idx, val, AL, AU, c, l, u, pairs, n, pair_probes, out);
  hipLaunchKernelGGL(k_scr_merge_rows, dim3(grid_for(row_probes)), dim3(kBlock), 0, s, ptr, idx, val, AL, AU, m,
                     pairs, row_probes, out + 1);
  return hipGetLastError();
}

hipError_t gpk_screen_structural_prefix(const int32_t* ptr, const int32_t* idx, const double* val,
                                        const double* AL, const double* AU, int32_t m, int32_t* out, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scr_structural_prefix, dim3(1), dim3(64), 0, s, ptr, idx, val, AL, AU, m, out);
  return hipGetLastError();
}

}  // extern "C"
