// Row-side reduction kernels: empty rows, singleton rows, activity-redundant
// rows, implied (dominated) bound tightening, FixedCol tape packing and
// duplicate rows.  Every floating-point reduction runs in one fixed order
// (CSR order along a row, AT order along a column) so results do not depend
// on scheduling.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>

#include "gp_device.hpp"
#include "gp_kernels.h"

using namespace gpdev;

namespace {

// ------------------------------------------------------------ generic ----
__global__ void k_row_shift(gpk_csr A, const uint8_t* keep_row, const uint8_t* col_flag, const double* col_val,
                            double abs_min, double* AL, double* AU) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.rows) return;
  double shift = 0.0;
  if (keep_row == nullptr || keep_row[i] != 0) {
    for (int32_t p = A.ptr[i]; p < A.ptr[i + 1]; ++p) {
      const int32_t j = A.idx[p];
      if (!col_flag[j]) continue;
      const double a = A.val[p];
      if (abs_min >= 0.0 && !(fabs(a) > abs_min)) continue;
      shift += a * col_val[j];
    }
  }
  AL[i] -= shift;
  AU[i] -= shift;
}

__global__ void k_pack_f64(const uint8_t* mask, const double* vals, const int32_t* prefix, int32_t n, double* out) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && mask[i]) out[prefix[i]] = vals[i];
}

__global__ void k_pack_index(const uint8_t* mask, const int32_t* prefix, int32_t n, int32_t* out) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && mask[i]) out[prefix[i]] = i;
}

// ------------------------------------------------------- empty rows ----
// flags[0]: live empty row infeasible at 0; flags[1]: removable empty row.
__global__ void k_plan_init(int32_t m, int32_t n, gpk_plan p, const double* c, const double* l, const double* u,
                            const double* AL, const double* AU) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    p.keep_col[i] = 1;
    p.c[i] = c[i];
    p.l[i] = l[i];
    p.u[i] = u[i];
  }
  if (i < m) {
    p.keep_row[i] = 1;
    p.AL[i] = AL[i];
    p.AU[i] = AU[i];
  }
}

__global__ void k_phase_init(int32_t m, int32_t n, gpk_plan p, const double* c, const double* l, const double* u,
                             const double* AL, const double* AU, gpk_csr A, gpk_csr AT, gpk_phase_stats st) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    p.keep_col[i] = 1;
    p.c[i] = c[i];
    p.l[i] = l[i];
    p.u[i] = u[i];
    const int32_t b = AT.ptr[i];
    const int32_t cnt = AT.ptr[i + 1] - b;
    st.col_nnz[i] = cnt;
    st.empty_col_mask[i] = cnt == 0 ? 1 : 0;
    st.column_singleton_mask[i] = cnt == 1 ? 1 : 0;
    st.column_singleton_row[i] = cnt == 1 ? AT.idx[b] : -1;
    st.column_singleton_val[i] = cnt == 1 ? AT.val[b] : 0.0;
  }
  if (i < m) {
    p.keep_row[i] = 1;
    p.AL[i] = AL[i];
    p.AU[i] = AU[i];
    const int32_t b = A.ptr[i];
    const int32_t cnt = A.ptr[i + 1] - b;
    st.row_nnz[i] = cnt;
    st.singleton_row_mask[i] = cnt == 1 ? 1 : 0;
    st.singleton_row_col[i] = cnt == 1 ? A.idx[b] : -1;
    st.singleton_row_val[i] = cnt == 1 ? A.val[b] : 0.0;
  }
}

__global__ void k_empty_rows(int32_t m, gpk_plan p, const int32_t* row_nnz, double tol, int32_t apply,
                             int32_t* flags) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !p.keep_row[i] || row_nnz[i] != 0) return;
  if (p.AL[i] <= tol && p.AU[i] >= -tol) {
    if (apply) p.keep_row[i] = 0; else wave_set_flag(&flags[1]);
  } else if (!apply) {
    wave_set_flag(&flags[0]);
  }
}

// --------------------------------------------------- singleton rows ----
__global__ void k_singleton_rows_bounds(int32_t m, int32_t n, gpk_plan p, const uint8_t* mask, const int32_t* col,
                                        const double* val, double ztol, double* cl, double* cu, uint8_t* remove) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  remove[i] = 0;
  if (!p.keep_row[i] || !mask[i]) return;
  const int32_t j = col[i];
  const double a = val[i];
  if (j < 0 || j >= n || !(fabs(a) > ztol)) return;
  const double lo = p.AL[i], hi = p.AU[i];
  atomic_max_f64(&cl[j], a > 0.0 ? lo / a : hi / a);
  atomic_min_f64(&cu[j], a > 0.0 ? hi / a : lo / a);
  remove[i] = 1;
}

// flags[0] infeasible, flags[1] some row removed, flags[2] some bound changed
__global__ void k_singleton_rows_check(int32_t m, int32_t n, gpk_plan p, const double* cl, const double* cu,
                                       const uint8_t* remove, double tol, int32_t* flags) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    if (cl[t] > cu[t] + tol) wave_set_flag(&flags[0]);
    if (cl[t] != p.l[t] || cu[t] != p.u[t]) wave_set_flag(&flags[2]);
  }
  if (t < m && remove[t]) wave_set_flag(&flags[1]);
}

__global__ void k_singleton_rows_commit(int32_t m, int32_t n, gpk_plan p, const double* cl, const double* cu,
                                        const uint8_t* remove) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m && remove[t]) p.keep_row[t] = 0;
  if (t < n) {
    p.l[t] = cl[t];
    p.u[t] = cu[t];
  }
}

// -------------------------------------------------- redundant rows ----
// flags[0] infeasible, flags[1] any action.
__global__ void k_redundant_rows_classify(gpk_csr A, gpk_plan p, const int32_t* row_nnz, double tol, uint8_t* full,
                                          uint8_t* dlo, uint8_t* dhi, int32_t* flags) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.rows) return;
  full[i] = 0;
  dlo[i] = 0;
  dhi[i] = 0;
  if (!p.keep_row[i] || row_nnz[i] <= 1) return;
  const double lo = p.AL[i], hi = p.AU[i];
  const bool lf = isfinite(lo), hf = isfinite(hi);
  if (lf && hf && fabs(hi - lo) <= tol) return;
  double rmin = 0.0, rmax = 0.0;
  for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
    const int32_t j = A.idx[q];
    const double a = A.val[q];
    if (a >= 0.0) {
      rmin += a * p.l[j];
      rmax += a * p.u[j];
    } else {
      rmin += a * p.u[j];
      rmax += a * p.l[j];
    }
  }
  if ((lf && rmax < lo - tol) || (hf && rmin > hi + tol)) {
    wave_set_flag(&flags[0]);
    return;
  }
  const bool lo_imp = !lf || rmin >= lo - tol;
  const bool hi_imp = !hf || rmax <= hi + tol;
  if (lo_imp && hi_imp) {
    full[i] = 1;
    wave_set_flag(&flags[1]);
  } else if (lf && lo_imp) {
    dlo[i] = 1;
    wave_set_flag(&flags[1]);
  } else if (hf && hi_imp) {
    dhi[i] = 1;
    wave_set_flag(&flags[1]);
  }
}

__global__ void k_redundant_rows_apply(int32_t m, gpk_plan p, const uint8_t* full, const uint8_t* dlo,
                                       const uint8_t* dhi) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  if (full[i]) p.keep_row[i] = 0;
  if (dlo[i]) p.AL[i] = -INFINITY;
  if (dhi[i]) p.AU[i] = INFINITY;
}

// ------------------------------------------------- implied bounds ----
__device__ inline void term_interval(double a, double lj, double uj, double* tmin, double* tmax) {
  if (a >= 0.0) {
    *tmin = a * lj;
    *tmax = a * uj;
  } else {
    *tmin = a * uj;
    *tmax = a * lj;
  }
}
__device__ inline double resid_min(double s, int32_t ninf, double tmin) {
  if (isfinite(tmin)) return ninf == 0 ? s - tmin : -INFINITY;
  if (tmin < 0.0) return ninf > 1 ? -INFINITY : s;
  return s;
}
__device__ inline double resid_max(double s, int32_t pinf, double tmax) {
  if (isfinite(tmax)) return pinf == 0 ? s - tmax : INFINITY;
  if (tmax > 0.0) return pinf > 1 ? INFINITY : s;
  return s;
}
__device__ inline void implied_from_row(double a, double lo, double hi, double rs_min, double rs_max, double* il,
                                        double* iu) {
  *il = -INFINITY;
  *iu = INFINITY;
  if (a > 0.0) {
    if (isfinite(lo) && isfinite(rs_max)) *il = (lo - rs_max) / a;
    if (isfinite(hi) && isfinite(rs_min)) *iu = (hi - rs_min) / a;
  } else {
    if (isfinite(hi) && isfinite(rs_min)) *il = (hi - rs_min) / a;
    if (isfinite(lo) && isfinite(rs_max)) *iu = (lo - rs_max) / a;
  }
}

// Walks column j of AT and calls f(i, il, iu) for each entry on a live row
// (kept, more than one entry) with |a| > ztol, in AT order, where il / iu is
// the bound that row implies for the column (the column's own term taken
// out of the row's activity).  Entries go U at a time: their row indices and
// values load together, then every row quantity of the U rows, so a chunk
// costs two dependent loads instead of three per entry.
template <class F>
__device__ inline void implied_walk(const gpk_csr& AT, const gpk_implied_row* __restrict__ rows, double ztol,
                                    int32_t j, double ol, double ou, F&& f) {
  constexpr int U = 4;
  const int32_t qb = AT.ptr[j], qe = AT.ptr[j + 1];
  for (int32_t q0 = qb; q0 < qe; q0 += U) {
    int32_t ri[U];
    double av[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int32_t q = min(q0 + k, qe - 1);
      ri[k] = AT.idx[q];
      av[k] = AT.val[q];
    }
    gpk_implied_row r[U];
#pragma unroll
    for (int k = 0; k < U; ++k) r[k] = rows[ri[k]];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      if (q0 + k >= qe || !r[k].live) continue;
      const double a = av[k];
      if (fabs(a) <= ztol) continue;
      double tmin, tmax;
      term_interval(a, ol, ou, &tmin, &tmax);
      const double rs_min = resid_min(r[k].smin, r[k].ninf, tmin);
      const double rs_max = resid_max(r[k].smax, r[k].pinf, tmax);
      double il, iu;
      implied_from_row(a, r[k].lo, r[k].hi, rs_min, rs_max, &il, &iu);
      f(ri[k], il, iu);
    }
  }
}

// One record per row: the activity summary, the plan's row bounds and
// whether the row takes part (kept, more than one entry).
__global__ void k_implied_pack_rows(int32_t m, gpk_plan p, const int32_t* row_nnz, const double* rmin,
                                    const double* rmax, const int32_t* rni, const int32_t* rpi,
                                    gpk_implied_row* rows) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  gpk_implied_row r;
  r.smin = rmin[i];
  r.smax = rmax[i];
  r.lo = p.AL[i];
  r.hi = p.AU[i];
  r.ninf = rni[i];
  r.pinf = rpi[i];
  r.live = p.keep_row[i] != 0 && row_nnz[i] > 1 ? 1 : 0;
  for (int k = 0; k < 5; ++k) r.pad[k] = 0;
  rows[i] = r;
}

// Candidate bounds per column, pulled along AT (no atomics): the max (min)
// over the column's live rows of the implied lower (upper) bound, seeded with
// the current bound.  Same values as the reference's per-entry CAS push.
__global__ __launch_bounds__(kBlock) void k_implied_candidates(gpk_csr AT, gpk_plan p,
                                                               const gpk_implied_row* __restrict__ rows, double ztol,
                                                               double* cand_l, double* cand_u) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= AT.rows) return;
  const double ol = p.l[j], ou = p.u[j];
  double cl = ol, cu = ou;
  implied_walk(AT, rows, ztol, j, ol, ou, [&](int32_t, double il, double iu) {
    double nl = ol, nu = ou;
    if (isfinite(il)) nl = fmax(nl, il);
    if (isfinite(iu)) nu = fmin(nu, iu);
    if (nl > ol && !(cl >= nl)) cl = nl;
    if (nu < ou && !(cu <= nu)) cu = nu;
  });
  cand_l[j] = cl;
  cand_u[j] = cu;
}

// rule_implied_variable_bounds.cu:513-649; flags[0] infeasible, [1] tightened, [2] fixed.
__global__ void k_implied_finalize(int32_t n, gpk_plan p, const double* ref_l, const double* ref_u,
                                   const double* cand_l, const double* cand_u, const double* vmax, int32_t preserve,
                                   double ztol, double ftol, double* fl, double* fu, uint8_t* lch, uint8_t* uch,
                                   uint8_t* fixed, double* fixed_val, int32_t* flags) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double lj = p.l[j], uj = p.u[j];
  double cl = cand_l[j], cu = cand_u[j];
  const double vm = vmax[j];
  double nl = lj, nu = uj;
  uint8_t lc = 0, uc = 0, fx = 0;
  double fat = 0.0;
  double el = lj, eu = uj;
  if (isfinite(cl)) el = fmax(el, cl);
  if (isfinite(cu)) eu = fmin(eu, cu);
  if (el > eu + ftol) {
    wave_set_flag(&flags[0]);
    return;
  }
  const bool was_free = !isfinite(ref_l[j]) && !isfinite(ref_u[j]);
  const bool box = isfinite(cl) && isfinite(cu);
  const bool zero_cost = p.c != nullptr && fabs(p.c[j]) <= ztol;
  const bool nearly_fixed = box && (cu <= cl || (cu - cl) * vm <= ftol);
  if (preserve && was_free && box && zero_cost && !nearly_fixed) {
    cl = lj;
    cu = uj;
  }
  if (isfinite(cl) && cl > lj && isfinite(uj)) {
    if (cl >= uj + ftol) {
      wave_set_flag(&flags[0]);
      return;
    }
    if (cl >= uj || (uj - cl) * vm <= ftol) {
      fx = 1;
      fat = uj;
      nl = uj;
      nu = uj;
    }
  }
  if (!fx && isfinite(cl) && cl > lj) {
    const bool tighten = !isfinite(lj) || ((cl - lj > ftol * 1.0e4) && (cl - lj > 1.0e-2 * fabs(lj)));
    if (tighten) {
      if (cl != nearbyint(cl)) cl -= 0.5 * ftol * fabs(cl);
      nl = cl;
      lc = 1;
    }
  }
  if (!fx && isfinite(cu) && cu < uj && isfinite(nl)) {
    if (cu <= nl - ftol) {
      wave_set_flag(&flags[0]);
      return;
    }
    if (cu <= nl || (cu - nl) * vm <= ftol) {
      fx = 1;
      fat = nl;
      nu = nl;
    }
  }
  if (!fx && isfinite(cu) && cu < uj) {
    const bool tighten = !isfinite(uj) || ((uj - cu > ftol * 1.0e4) && (uj - cu > 1.0e-2 * fabs(uj)));
    if (tighten) {
      if (cu != nearbyint(cu)) cu += 0.5 * ftol * fabs(cu);
      nu = cu;
      uc = 1;
    }
  }
  if (fx) {
    lc = 0;
    uc = 0;
    wave_set_flag(&flags[2]);
  }
  if (lc || uc) wave_set_flag(&flags[1]);
  fl[j] = nl;
  fu[j] = nu;
  lch[j] = lc;
  uch[j] = uc;
  fixed[j] = fx;
  fixed_val[j] = fat;
}

// rule_implied_variable_bounds.cu:668-750 plus the per-column tape record count.
__global__ __launch_bounds__(kBlock) void k_implied_support(gpk_csr AT, gpk_plan p,
                                                            const gpk_implied_row* __restrict__ rows,
                                                            const uint8_t* lch, const uint8_t* uch, const double* fl,
                                                            const double* fu, double ztol, double tol, double stol,
                                                            int32_t* sup_l, int32_t* sup_u, int32_t* rec_count) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= AT.rows) return;
  int32_t sl = INT_MAX, su = INT_MAX;
  const bool need_l = lch[j] != 0, need_u = uch[j] != 0;
  if (need_l || need_u) {
    const double ol = p.l[j], ou = p.u[j];
    const double cl = fl[j], cu = fu[j];
    implied_walk(AT, rows, ztol, j, ol, ou, [&](int32_t i, double il, double iu) {
      if (need_l && cl > ol + tol && isfinite(il) && fabs(il - cl) <= stol) sl = min(sl, i);
      if (need_u && cu < ou - tol && isfinite(iu) && fabs(iu - cu) <= stol) su = min(su, i);
    });
  }
  sup_l[j] = sl;
  sup_u[j] = su;
  rec_count[j] = (need_l ? 1 : 0) + (need_u ? 1 : 0);
}

// rule_implied_variable_bounds.cu:771-804: records {col, row, old_l, old_u, new_l, new_u}.
__global__ void k_implied_pack_tape(int32_t n, gpk_plan p, const uint8_t* lch, const uint8_t* uch,
                                    const int32_t* sup_l, const int32_t* sup_u, const double* fl, const double* fu,
                                    const int32_t* prefix, int32_t* rcol, int32_t* rrow, double* rvals) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int32_t out = prefix[j];
  double cur_l = p.l[j];
  const double cur_u = p.u[j];
  if (lch[j]) {
    rcol[out] = j;
    rrow[out] = sup_l[j];
    rvals[4 * out + 0] = cur_l;
    rvals[4 * out + 1] = cur_u;
    rvals[4 * out + 2] = fl[j];
    rvals[4 * out + 3] = cur_u;
    cur_l = fl[j];
    ++out;
  }
  if (uch[j]) {
    rcol[out] = j;
    rrow[out] = sup_u[j];
    rvals[4 * out + 0] = cur_l;
    rvals[4 * out + 1] = cur_u;
    rvals[4 * out + 2] = cur_l;
    rvals[4 * out + 3] = fu[j];
  }
}

__global__ void k_implied_apply(int32_t n, gpk_plan p, const double* fl, const double* fu, const uint8_t* fixed) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  p.l[j] = fl[j];
  p.u[j] = fu[j];
  if (fixed[j]) p.keep_col[j] = 0;
}

// flags[3] infeasible empty row, flags[4] removed empty row.
__global__ void k_remove_empty_rows_after_fix(gpk_csr A, gpk_plan p, double ftol, int32_t* flags) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.rows || !p.keep_row[i]) return;
  int32_t cnt = 0;
  for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) cnt += p.keep_col[A.idx[q]] ? 1 : 0;
  if (cnt != 0) return;
  if (p.AL[i] <= ftol && p.AU[i] >= -ftol) {
    p.keep_row[i] = 0;
    wave_set_flag(&flags[4]);
  } else {
    wave_set_flag(&flags[3]);
  }
}

// ------------------------------------------------ FixedCol packing ----
__global__ void k_fixed_col_counts(gpk_csr AT, const uint8_t* fixed, const uint8_t* excluded,
                                   const uint8_t* keep_row, uint8_t* selected, int32_t* active) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= AT.rows) return;
  const bool sel = fixed[j] && (excluded == nullptr || !excluded[j]);
  selected[j] = sel ? 1 : 0;
  int32_t cnt = 0;
  if (sel)
    for (int32_t q = AT.ptr[j]; q < AT.ptr[j + 1]; ++q) cnt += keep_row[AT.idx[q]] ? 1 : 0;
  active[j] = cnt;
}

__global__ void k_fixed_col_pack(gpk_csr AT, const uint8_t* selected, const int32_t* sel_prefix,
                                 const int32_t* nnz_prefix, const double* fixed_val, const double* c,
                                 const uint8_t* keep_row, int32_t* out_col, double* out_v, double* out_c,
                                 int32_t* out_start, int32_t* out_rows, double* out_coef) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= AT.rows || !selected[j]) return;
  const int32_t r = sel_prefix[j];
  out_col[r] = j;
  out_start[r] = nnz_prefix[j];
  out_v[r] = fixed_val[j];
  out_c[r] = c ? c[j] : 0.0;
  int32_t o = nnz_prefix[j];
  for (int32_t q = AT.ptr[j]; q < AT.ptr[j + 1]; ++q) {
    if (!keep_row[AT.idx[q]]) continue;
    out_rows[o] = AT.idx[q];
    out_coef[o] = AT.val[q];
    ++o;
  }
}

// --------------------------------------------------- duplicate rows ----
constexpr uint64_t kFnvOffset = 0xcbf29ce484222325ULL;
constexpr uint64_t kFnvPrime = 0x100000001b3ULL;
__device__ inline uint64_t fnv_mix(uint64_t h, uint64_t x) { return (h ^ x) * kFnvPrime; }
__device__ inline uint64_t dbits(double v) { return static_cast<uint64_t>(__double_as_longlong(v)); }

// rule_duplicate_rows.cu:78-114 (sequential FNV over len, col, a/a_first).
__global__ void k_dup_row_hash(gpk_csr A, const uint8_t* keep_row, double ctol, uint64_t* hash) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.rows) return;
  if (!keep_row[i]) {
    hash[i] = static_cast<uint64_t>(i);
    return;
  }
  const int32_t b = A.ptr[i], e = A.ptr[i + 1];
  if (e - b <= 0 || fabs(A.val[b]) <= ctol) {
    hash[i] = fnv_mix(kFnvOffset, static_cast<uint64_t>(i));
    return;
  }
  const double piv = A.val[b];
  uint64_t h = fnv_mix(kFnvOffset, static_cast<uint64_t>(e - b));
  for (int32_t q = b; q < e; ++q) {
    h = fnv_mix(h, static_cast<uint64_t>(static_cast<int64_t>(A.idx[q])));
    h = fnv_mix(h, dbits(A.val[q] / piv));
  }
  hash[i] = h;
}

// rule_duplicate_rows.cu:116-178, the hash the reference uses when
// nnz > 32 m (:484-486): one FNV pair hash (col, a / a_first) per entry,
// XOR-combined (order-free, so 64 lanes give the value 32 do) and mixed into
// the length seed.  One wave per row.
__global__ void k_dup_row_hash_wide(gpk_csr A, const uint8_t* keep_row, double ctol, uint64_t* hash) {
  const int32_t i = static_cast<int32_t>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = lane_id();
  if (i >= A.rows) return;
  if (!keep_row[i]) {
    if (lane == 0) hash[i] = static_cast<uint64_t>(i);
    return;
  }
  const int32_t b = A.ptr[i], e = A.ptr[i + 1];
  if (e - b <= 0 || fabs(A.val[b]) <= ctol) {
    if (lane == 0) hash[i] = fnv_mix(kFnvOffset, static_cast<uint64_t>(i));
    return;
  }
  const double piv = A.val[b];
  uint64_t agg = 0;
  for (int32_t q = b + lane; q < e; q += 64)
    agg ^= fnv_mix(fnv_mix(kFnvOffset, static_cast<uint64_t>(static_cast<int64_t>(A.idx[q]))), dbits(A.val[q] / piv));
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) agg ^= static_cast<uint64_t>(__shfl_xor(static_cast<unsigned long long>(agg), d, 64));
  if (lane == 0) hash[i] = fnv_mix(fnv_mix(kFnvOffset, static_cast<uint64_t>(e - b)), agg);
}

__device__ bool dup_row_ratio(const gpk_csr& A, int32_t i, int32_t k, double ctol, double* r_out) {
  const int32_t bi = A.ptr[i], li = A.ptr[i + 1] - bi;
  const int32_t bk = A.ptr[k], lk = A.ptr[k + 1] - bk;
  if (li != lk) return false;
  if (li <= 0) {
    *r_out = 1.0;
    return true;
  }
  double r = 0.0;
  bool set = false;
  for (int32_t o = 0; o < li; ++o) {
    if (A.idx[bi + o] != A.idx[bk + o]) return false;
    const double ai = A.val[bi + o], ak = A.val[bk + o];
    if (!set) {
      if (fabs(ak) <= ctol) return false;
      r = ai / ak;
      set = true;
    }
    if (fabs(ai - r * ak) > ctol) return false;
  }
  *r_out = r;
  return set;
}

__device__ inline void scaled_interval(double lo, double hi, double r, double* sl, double* su) {
  const double a = lo * r, b = hi * r;
  if (r >= 0.0) {
    *sl = a;
    *su = b;
  } else {
    *sl = b;
    *su = a;
  }
}

// rule_duplicate_rows.cu:243-381, one thread per hash bucket (its first slot).
// rec_vals[5*row..]: ratio, kept_old_AL, kept_old_AU, deleted_old_AL, deleted_old_AU.
__global__ void k_dup_row_groups(gpk_csr A, gpk_plan p, const uint64_t* sh, const int32_t* srow, double ctol,
                                 double itol, uint8_t* row_delete, int32_t* merge_to, int32_t* merge_step,
                                 double* rec, int32_t* flags) {
  const int32_t m = A.rows;
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  const uint64_t hs = sh[s];
  if (s > 0 && sh[s - 1] == hs) return;
  int32_t e = s;
  while (e + 1 < m && sh[e + 1] == hs) ++e;
  if (e <= s) return;
  int32_t seed = -1;
  bool seed_eq = false;
  for (int32_t a = s; a <= e; ++a) {
    const int32_t row = srow[a];
    if (!p.keep_row[row]) continue;
    const double lo = p.AL[row], hi = p.AU[row];
    const bool eq = isfinite(lo) && isfinite(hi) && fabs(lo - hi) <= itol;
    if (seed < 0 || (eq && !seed_eq)) {
      seed = row;
      seed_eq = eq;
    }
  }
  if (seed < 0) return;
  int32_t rep = seed;
  double rsl = p.AL[seed], rsu = p.AU[seed];
  for (int32_t a = s; a <= e; ++a) {
    const int32_t row = srow[a];
    if (row == seed || !p.keep_row[row]) continue;
    double r;
    if (!dup_row_ratio(A, seed, row, ctol, &r)) continue;
    double sl, su;
    scaled_interval(p.AL[row], p.AU[row], r, &sl, &su);
    if (rsl <= sl + itol && su <= rsu + itol) {
      rep = row;
      rsl = sl;
      rsu = su;
    }
  }
  double rl = p.AL[rep], ru = p.AU[rep];
  bool peer = false;
  int32_t step = 0;
  for (int32_t a = s; a <= e; ++a) {
    const int32_t row = srow[a];
    if (row == rep || !p.keep_row[row]) continue;
    double r;
    if (!dup_row_ratio(A, rep, row, ctol, &r)) continue;
    peer = true;
    double sl, su;
    scaled_interval(p.AL[row], p.AU[row], r, &sl, &su);
    if (fmax(rl, sl) > fmin(ru, su) + itol) {
      wave_set_flag(&flags[0]);
      return;
    }
    ++step;
    merge_to[row] = rep;
    merge_step[row] = step;
    rec[5 * static_cast<int64_t>(row) + 0] = r;
    rec[5 * static_cast<int64_t>(row) + 1] = rl;
    rec[5 * static_cast<int64_t>(row) + 2] = ru;
    rec[5 * static_cast<int64_t>(row) + 3] = p.AL[row];
    rec[5 * static_cast<int64_t>(row) + 4] = p.AU[row];
    rl = fmax(rl, sl);
    ru = fmin(ru, su);
    row_delete[row] = 1;
  }
  if (peer) {
    p.AL[rep] = rl;
    p.AU[rep] = ru;
  }
}

__global__ void k_delete_rows(int32_t m, uint8_t* keep_row, const uint8_t* row_delete, int32_t* deleted) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && row_delete[i]) {
    keep_row[i] = 0;
    wave_set_flag(deleted);
  }
}

// Deleted duplicate rows packed in ascending row order (prefix = exclusive
// scan of del): row index, kept partner, merge step and the five record
// values, so only the deleted rows' data crosses to the host.
__global__ void k_dup_row_pack(int32_t m, const uint8_t* del, const int32_t* prefix, const int32_t* merge_to,
                               const int32_t* step, const double* rec, int32_t* out_row, int32_t* out_to,
                               int32_t* out_step, double* out_rec) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !del[i]) return;
  const int32_t r = prefix[i];
  out_row[r] = i;
  out_to[r] = merge_to[i];
  out_step[r] = step[i];
  for (int t = 0; t < 5; ++t) out_rec[5 * static_cast<int64_t>(r) + t] = rec[5 * static_cast<int64_t>(i) + t];
}

// ---- batched singleton closure (gpu_presolve.cu:98-247) ---------------------
// One launch over max(m, n) threads: thread t takes column t and row t.
// Live degrees lose the entries of columns (rows) removed since the previous
// refresh; integer atomics, so the result does not depend on their order.
// Each thread then copies its own entry of the current masks into the
// "before" masks (no other thread reads that entry).
__global__ void k_closure_degrees(gpk_csr A, gpk_csr AT, const uint8_t* keep_row, const uint8_t* keep_col,
                                  uint8_t* row_before, uint8_t* col_before, int32_t* row_nnz, int32_t* col_nnz) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < A.cols) {
    const uint8_t kc = keep_col[t];
    if (col_before[t] != 0 && kc == 0) {
      for (int32_t q = AT.ptr[t]; q < AT.ptr[t + 1]; ++q) {
        const int32_t row = AT.idx[q];
        if (keep_row[row] != 0) atomicSub(&row_nnz[row], 1);
      }
    }
    col_before[t] = kc;
  }
  if (t < A.rows) {
    const uint8_t kr = keep_row[t];
    if (row_before[t] != 0 && kr == 0) {
      for (int32_t q = A.ptr[t]; q < A.ptr[t + 1]; ++q) {
        const int32_t col = A.idx[q];
        if (keep_col[col] != 0) atomicSub(&col_nnz[col], 1);
      }
    }
    row_before[t] = kr;
  }
}

// Singleton rows and columns from the live degrees; the singleton entry is
// the first live one of the row (column).
__global__ void k_closure_singletons(gpk_csr A, gpk_csr AT, const uint8_t* keep_row, const uint8_t* keep_col,
                                     const int32_t* row_nnz, const int32_t* col_nnz, uint8_t* srow_mask,
                                     int32_t* srow_col, double* srow_val, uint8_t* scol_mask, int32_t* scol_row,
                                     double* scol_val) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < A.rows) {
    const bool single = keep_row[t] != 0 && row_nnz[t] == 1;
    int32_t col = -1;
    double val = 0.0;
    if (single) {
      for (int32_t q = A.ptr[t]; q < A.ptr[t + 1]; ++q) {
        if (keep_col[A.idx[q]] == 0) continue;
        col = A.idx[q];
        val = A.val[q];
        break;
      }
    }
    srow_mask[t] = single ? 1 : 0;
    srow_col[t] = col;
    srow_val[t] = val;
  }
  if (t < A.cols) {
    const bool single = keep_col[t] != 0 && col_nnz[t] == 1;
    int32_t row = -1;
    double val = 0.0;
    if (single) {
      for (int32_t q = AT.ptr[t]; q < AT.ptr[t + 1]; ++q) {
        if (keep_row[AT.idx[q]] == 0) continue;
        row = AT.idx[q];
        val = AT.val[q];
        break;
      }
    }
    scol_mask[t] = single ? 1 : 0;
    scol_row[t] = row;
    scol_val[t] = val;
  }
}

// Snapshot of the plan state the dual-fix candidates are measured against.
__global__ void k_closure_snapshot(int32_t m, int32_t n, gpk_plan plan, gpk_plan seen) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    seen.c[t] = plan.c[t];
    seen.l[t] = plan.l[t];
    seen.u[t] = plan.u[t];
  }
  if (t < m) {
    seen.keep_row[t] = plan.keep_row[t];
    seen.AL[t] = plan.AL[t];
    seen.AU[t] = plan.AU[t];
  }
}

// Live columns whose c, l or u changed since the snapshot.
__global__ void k_closure_mark_cols(int32_t n, gpk_plan plan, gpk_plan seen, uint8_t* cand) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  cand[j] = plan.keep_col[j] != 0 && (plan.c[j] != seen.c[j] || plan.l[j] != seen.l[j] || plan.u[j] != seen.u[j])
                ? 1
                : 0;
}

// Live columns of rows removed, or with AL/AU changed, since the snapshot.
__global__ void k_closure_mark_rows(gpk_csr A, gpk_plan plan, gpk_plan seen, uint8_t* cand) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.rows) return;
  if (plan.keep_row[i] == seen.keep_row[i] && plan.AL[i] == seen.AL[i] && plan.AU[i] == seen.AU[i]) return;
  for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
    const int32_t col = A.idx[q];
    if (plan.keep_col[col] != 0) cand[col] = 1;
  }
}

}  // namespace

#define GPK_LAUNCH(kernel, n, ...)                                                      \
  do {                                                                                  \
    if ((n) > 0) hipLaunchKernelGGL(kernel, dim3(grid_for(n)), dim3(kBlock), 0, stream, __VA_ARGS__); \
    return hipGetLastError();                                                           \
  } while (0)

extern "C" {

hipError_t gpk_row_shift(gpk_csr A, const uint8_t* keep_row, const uint8_t* col_flag, const double* col_val,
                         double abs_min, double* AL, double* AU, hipStream_t stream) {
  GPK_LAUNCH(k_row_shift, A.rows, A, keep_row, col_flag, col_val, abs_min, AL, AU);
}
hipError_t gpk_pack_f64(const uint8_t* mask, const double* vals, const int32_t* prefix, int32_t n, double* out,
                        hipStream_t stream) {
  GPK_LAUNCH(k_pack_f64, n, mask, vals, prefix, n, out);
}
hipError_t gpk_pack_index(const uint8_t* mask, const int32_t* prefix, int32_t n, int32_t* out, hipStream_t stream) {
  GPK_LAUNCH(k_pack_index, n, mask, prefix, n, out);
}
hipError_t gpk_phase_init(int32_t m, int32_t n, gpk_plan plan, const double* c, const double* l, const double* u,
                          const double* AL, const double* AU, gpk_csr A, gpk_csr AT, gpk_phase_stats stats,
                          hipStream_t stream) {
  GPK_LAUNCH(k_phase_init, m > n ? m : n, m, n, plan, c, l, u, AL, AU, A, AT, stats);
}

hipError_t gpk_plan_init(int32_t m, int32_t n, gpk_plan plan, const double* c, const double* l, const double* u,
                         const double* AL, const double* AU, hipStream_t stream) {
  GPK_LAUNCH(k_plan_init, m > n ? m : n, m, n, plan, c, l, u, AL, AU);
}
hipError_t gpk_empty_rows(int32_t m, gpk_plan plan, const int32_t* row_nnz, double tol, int32_t apply, int32_t* flags,
                          hipStream_t stream) {
  GPK_LAUNCH(k_empty_rows, m, m, plan, row_nnz, tol, apply, flags);
}
hipError_t gpk_singleton_rows_bounds(int32_t m, int32_t n, gpk_plan plan, const uint8_t* mask, const int32_t* col,
                                     const double* val, double zero_tol, double* cand_l, double* cand_u,
                                     uint8_t* row_remove, hipStream_t stream) {
  GPK_LAUNCH(k_singleton_rows_bounds, m, m, n, plan, mask, col, val, zero_tol, cand_l, cand_u, row_remove);
}
hipError_t gpk_singleton_rows_check(int32_t m, int32_t n, gpk_plan plan, const double* cand_l, const double* cand_u,
                                    const uint8_t* row_remove, double bound_tol, int32_t* flags, hipStream_t stream) {
  const int32_t mn = m > n ? m : n;
  GPK_LAUNCH(k_singleton_rows_check, mn, m, n, plan, cand_l, cand_u, row_remove, bound_tol, flags);
}
hipError_t gpk_singleton_rows_commit(int32_t m, int32_t n, gpk_plan plan, const double* cand_l, const double* cand_u,
                                     const uint8_t* row_remove, hipStream_t stream) {
  const int32_t mn = m > n ? m : n;
  GPK_LAUNCH(k_singleton_rows_commit, mn, m, n, plan, cand_l, cand_u, row_remove);
}
hipError_t gpk_redundant_rows_classify(gpk_csr A, gpk_plan plan, const int32_t* row_nnz, double tol, uint8_t* full,
                                       uint8_t* dlo, uint8_t* dhi, int32_t* flags, hipStream_t stream) {
  GPK_LAUNCH(k_redundant_rows_classify, A.rows, A, plan, row_nnz, tol, full, dlo, dhi, flags);
}
hipError_t gpk_redundant_rows_apply(int32_t m, gpk_plan plan, const uint8_t* full, const uint8_t* dlo,
                                    const uint8_t* dhi, hipStream_t stream) {
  GPK_LAUNCH(k_redundant_rows_apply, m, m, plan, full, dlo, dhi);
}
hipError_t gpk_implied_pack_rows(int32_t m, gpk_plan plan, const int32_t* row_nnz, const double* rmin,
                                 const double* rmax, const int32_t* rni, const int32_t* rpi, gpk_implied_row* rows,
                                 hipStream_t stream) {
  GPK_LAUNCH(k_implied_pack_rows, m, m, plan, row_nnz, rmin, rmax, rni, rpi, rows);
}
hipError_t gpk_implied_candidates(gpk_csr AT, gpk_plan plan, const gpk_implied_row* rows, double zero_tol,
                                  double* cand_l, double* cand_u, hipStream_t stream) {
  GPK_LAUNCH(k_implied_candidates, AT.rows, AT, plan, rows, zero_tol, cand_l, cand_u);
}
hipError_t gpk_implied_finalize(int32_t n, gpk_plan plan, const double* ref_l, const double* ref_u,
                                const double* cand_l, const double* cand_u, const double* vmax,
                                int32_t preserve_free_zero_cost, double zero_tol, double feas_tol, double* fin_l,
                                double* fin_u, uint8_t* lch, uint8_t* uch, uint8_t* fixed, double* fixed_val,
                                int32_t* flags, hipStream_t stream) {
  GPK_LAUNCH(k_implied_finalize, n, n, plan, ref_l, ref_u, cand_l, cand_u, vmax, preserve_free_zero_cost, zero_tol,
             feas_tol, fin_l, fin_u, lch, uch, fixed, fixed_val, flags);
}
hipError_t gpk_implied_support(gpk_csr AT, gpk_plan plan, const gpk_implied_row* rows, const uint8_t* lch,
                               const uint8_t* uch, const double* fin_l, const double* fin_u, double zero_tol,
                               double tol, double support_tol, int32_t* support_l, int32_t* support_u,
                               int32_t* rec_count, hipStream_t stream) {
  GPK_LAUNCH(k_implied_support, AT.rows, AT, plan, rows, lch, uch, fin_l, fin_u, zero_tol, tol, support_tol,
             support_l, support_u, rec_count);
}
hipError_t gpk_implied_pack_tape(int32_t n, gpk_plan plan, const uint8_t* lch, const uint8_t* uch,
                                 const int32_t* support_l, const int32_t* support_u, const double* fin_l,
                                 const double* fin_u, const int32_t* prefix, int32_t* rec_col, int32_t* rec_row,
                                 double* rec_vals, hipStream_t stream) {
  GPK_LAUNCH(k_implied_pack_tape, n, n, plan, lch, uch, support_l, support_u, fin_l, fin_u, prefix, rec_col, rec_row,
             rec_vals);
}
hipError_t gpk_implied_apply(int32_t n, gpk_plan plan, const double* fin_l, const double* fin_u, const uint8_t* fixed,
                             hipStream_t stream) {
  GPK_LAUNCH(k_implied_apply, n, n, plan, fin_l, fin_u, fixed);
}
hipError_t gpk_remove_empty_rows_after_fix(gpk_csr A, gpk_plan plan, double feas_tol, int32_t* flags,
                                           hipStream_t stream) {
  GPK_LAUNCH(k_remove_empty_rows_after_fix, A.rows, A, plan, feas_tol, flags);
}
hipError_t gpk_fixed_col_counts(gpk_csr AT, const uint8_t* fixed, const uint8_t* excluded, const uint8_t* keep_row,
                                uint8_t* selected, int32_t* active_nnz, hipStream_t stream) {
  GPK_LAUNCH(k_fixed_col_counts, AT.rows, AT, fixed, excluded, keep_row, selected, active_nnz);
}
hipError_t gpk_fixed_col_pack(gpk_csr AT, const uint8_t* selected, const int32_t* sel_prefix,
                              const int32_t* nnz_prefix, const double* fixed_val, const double* c,
                              const uint8_t* keep_row, int32_t* out_col, double* out_v, double* out_c,
                              int32_t* out_start, int32_t* out_rows, double* out_coef, hipStream_t stream) {
  GPK_LAUNCH(k_fixed_col_pack, AT.rows, AT, selected, sel_prefix, nnz_prefix, fixed_val, c, keep_row, out_col, out_v,
             out_c, out_start, out_rows, out_coef);
}
hipError_t gpk_dup_row_hash(gpk_csr A, const uint8_t* keep_row, double coeff_tol, uint64_t* hash,
                            hipStream_t stream) {
  if (static_cast<int64_t>(A.nnz) > 32LL * static_cast<int64_t>(A.rows)) {
    if (A.rows > 0)
      hipLaunchKernelGGL(k_dup_row_hash_wide, dim3(grid_for(static_cast<int64_t>(A.rows) * 64)), dim3(kBlock), 0,
                         stream, A, keep_row, coeff_tol, hash);
    return hipGetLastError();
  }
  GPK_LAUNCH(k_dup_row_hash, A.rows, A, keep_row, coeff_tol, hash);
}
hipError_t gpk_dup_row_groups(gpk_csr A, gpk_plan plan, const uint64_t* sorted_hash, const int32_t* sorted_rows,
                              double coeff_tol, double interval_tol, uint8_t* row_delete, int32_t* merge_to,
                              int32_t* merge_step, double* rec_vals, int32_t* flags, hipStream_t stream) {
  GPK_LAUNCH(k_dup_row_groups, A.rows, A, plan, sorted_hash, sorted_rows, coeff_tol, interval_tol, row_delete,
             merge_to, merge_step, rec_vals, flags);
}
hipError_t gpk_delete_rows(int32_t m, uint8_t* keep_row, const uint8_t* row_delete, int32_t* deleted,
                           hipStream_t stream) {
  GPK_LAUNCH(k_delete_rows, m, m, keep_row, row_delete, deleted);
}
hipError_t gpk_dup_row_pack(int32_t m, const uint8_t* del, const int32_t* prefix, const int32_t* merge_to,
                            const int32_t* step, const double* rec, int32_t* out_row, int32_t* out_to,
                            int32_t* out_step, double* out_rec, hipStream_t stream) {
  GPK_LAUNCH(k_dup_row_pack, m, m, del, prefix, merge_to, step, rec, out_row, out_to, out_step, out_rec);
}

hipError_t gpk_closure_degrees(gpk_csr A, gpk_csr AT, const uint8_t* keep_row, const uint8_t* keep_col,
                               uint8_t* row_before, uint8_t* col_before, int32_t* row_nnz, int32_t* col_nnz,
                               hipStream_t stream) {
  GPK_LAUNCH(k_closure_degrees, A.rows > A.cols ? A.rows : A.cols, A, AT, keep_row, keep_col, row_before, col_before,
             row_nnz, col_nnz);
}
hipError_t gpk_closure_singletons(gpk_csr A, gpk_csr AT, const uint8_t* keep_row, const uint8_t* keep_col,
                                  const int32_t* row_nnz, const int32_t* col_nnz, uint8_t* srow_mask,
                                  int32_t* srow_col, double* srow_val, uint8_t* scol_mask, int32_t* scol_row,
                                  double* scol_val, hipStream_t stream) {
  GPK_LAUNCH(k_closure_singletons, A.rows > A.cols ? A.rows : A.cols, A, AT, keep_row, keep_col, row_nnz, col_nnz,
             srow_mask, srow_col, srow_val, scol_mask, scol_row, scol_val);
}
hipError_t gpk_closure_snapshot(int32_t m, int32_t n, gpk_plan plan, gpk_plan seen, hipStream_t stream) {
  GPK_LAUNCH(k_closure_snapshot, m > n ? m : n, m, n, plan, seen);
}
hipError_t gpk_closure_candidates(gpk_csr A, gpk_plan plan, gpk_plan seen, uint8_t* cand, hipStream_t stream) {
  if (A.cols > 0) hipLaunchKernelGGL(k_closure_mark_cols, dim3(grid_for(A.cols)), dim3(kBlock), 0, stream, A.cols,
                                     plan, seen, cand);
  if (A.rows > 0) hipLaunchKernelGGL(k_closure_mark_rows, dim3(grid_for(A.rows)), dim3(kBlock), 0, stream, A, plan,
                                     seen, cand);
  return hipGetLastError();
}

}  // extern "C"
