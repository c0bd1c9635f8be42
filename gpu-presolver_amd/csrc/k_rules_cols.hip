// Column-side reduction kernels: fixed columns, empty columns, column
// singletons, dual fixing, duplicate columns, doubleton equations (batch)
// and redundant bounds.  Floating-point reductions run in one fixed order
// (CSR order along a row, AT order along a column).
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>

#include "gp_device.hpp"
#include "gp_kernels.h"

using namespace gpdev;

namespace {

// ---------------------------------------------------- fixed / empty ----
__global__ void k_fixed_vars_candidates(int32_t n, gpk_plan p, double tol, uint8_t* fixed, double* v,
                                        double* contrib, int32_t* count) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  fixed[j] = 0;
  v[j] = 0.0;
  contrib[j] = 0.0;
  if (!p.keep_col[j]) return;
  const double lj = p.l[j], uj = p.u[j];
  if (isfinite(lj) && isfinite(uj) && fabs(uj - lj) <= tol) {
    const double vj = lj + 0.5 * (uj - lj);
    fixed[j] = 1;
    v[j] = vj;
    contrib[j] = p.c[j] * vj;
    wave_set_flag(count);
  }
}

// flags[0] infeasible, flags[1] unbounded, flags[2] any fixed.
__global__ void k_empty_cols(int32_t n, gpk_plan p, const uint8_t* empty, double btol, double ztol, uint8_t* fixed,
                             double* v, double* contrib, int32_t* flags) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  fixed[j] = 0;
  v[j] = 0.0;
  contrib[j] = 0.0;
  if (!(p.keep_col[j] && empty[j])) return;
  const double cj = p.c[j], lj = p.l[j], uj = p.u[j];
  if (lj > uj + btol) {
    wave_set_flag(&flags[0]);
    return;
  }
  if ((cj > ztol && !isfinite(lj)) || (cj < -ztol && !isfinite(uj))) {
    wave_set_flag(&flags[1]);
    return;
  }
  double val = 0.0;
  if (cj > ztol) val = lj;
  else if (cj < -ztol) val = uj;
  else if (isfinite(lj)) val = lj;
  else if (isfinite(uj)) val = uj;
  fixed[j] = 1;
  v[j] = val;
  contrib[j] = cj * val;
  wave_set_flag(&flags[2]);
}

__global__ void k_fix_cols_apply(int32_t n, gpk_plan p, const uint8_t* fixed, const double* v) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || !fixed[j]) return;
  p.l[j] = v[j];
  p.u[j] = v[j];
  p.keep_col[j] = 0;
}

// ------------------------------------------------ column singletons ----
constexpr uint8_t kIneqNone = 0, kIneqElim = 1, kIneqLhsToRhs = 2, kIneqRhsToLhs = 3;

__device__ inline bool cs_direct_unbounded(double cj, double a, double lhs, double rhs, double lb, double ub,
                                           double zt) {
  return (cj > zt && a > zt && !isfinite(lhs) && !isfinite(lb)) ||
         (cj > zt && a < -zt && !isfinite(rhs) && !isfinite(lb)) ||
         (cj < -zt && a < -zt && !isfinite(lhs) && !isfinite(ub)) ||
         (cj < -zt && a > zt && !isfinite(rhs) && !isfinite(ub));
}
__device__ inline bool cs_free_above(double a, double lhs, double rhs, double ub, double rmin, double rmax,
                                     double tol) {
  if (!isfinite(ub)) return true;
  double iu = INFINITY;
  if (a > 0.0 && isfinite(rhs)) iu = (rhs - rmin) / a;
  else if (a < 0.0 && isfinite(lhs)) iu = (lhs - rmax) / a;
  return iu <= ub + tol;
}
__device__ inline bool cs_free_below(double a, double lhs, double rhs, double lb, double rmin, double rmax,
                                     double tol) {
  if (!isfinite(lb)) return true;
  double il = -INFINITY;
  if (a > 0.0 && isfinite(lhs)) il = (lhs - rmax) / a;
  else if (a < 0.0 && isfinite(rhs)) il = (rhs - rmin) / a;
  return il >= lb - tol;
}
__device__ inline double cs_active_side(double cj, double a, double lhs, double rhs, double zt) {
  if ((cj > zt && a > 0.0) || (cj < -zt && a < 0.0)) return lhs;
  if ((cj > zt && a < 0.0) || (cj < -zt && a > 0.0)) return rhs;
  return isfinite(lhs) ? lhs : rhs;
}
__device__ inline uint8_t cs_ineq_action(double cj, double a, double lhs, double rhs, bool above, bool below,
                                         double zt, double* side) {
  if (above && below) {
    *side = cs_active_side(cj, a, lhs, rhs, zt);
    return isfinite(*side) ? kIneqElim : kIneqNone;
  }
  if (((cj < -zt && a > 0.0 && above) || (cj > zt && a < 0.0 && below)) && isfinite(rhs)) {
    *side = rhs;
    return kIneqLhsToRhs;
  }
  if (((cj > zt && a > 0.0 && below) || (cj < -zt && a < 0.0 && above)) && isfinite(lhs)) {
    *side = lhs;
    return kIneqRhsToLhs;
  }
  *side = 0.0;
  return kIneqNone;
}

// rule_column_singletons.cu:653-713: activity of the rest of the row.
__device__ bool cs_rest_bounds(const gpk_row_cache& rc, int32_t row, int32_t col, double a, const double* l,
                               const double* u, double* rest_min, double* rest_max) {
  if (rc.invalid[row]) return false;
  const double tmin = a >= 0.0 ? a * l[col] : a * u[col];
  const double tmax = a >= 0.0 ? a * u[col] : a * l[col];
  if (isnan(tmin) || isnan(tmax)) return false;
  int32_t mn = rc.mn[row], mp = rc.mp[row], xn = rc.xn[row], xp = rc.xp[row];
  double vmin = rc.fmin[row], vmax = rc.fmax[row];
  if (isinf(tmin)) {
    if (tmin < 0.0) --mn; else --mp;
  } else {
    vmin -= tmin;
  }
  if (isinf(tmax)) {
    if (tmax < 0.0) --xn; else --xp;
  } else {
    vmax -= tmax;
  }
  if (mn < 0 || mp < 0 || xn < 0 || xp < 0 || (mn > 0 && mp > 0) || (xn > 0 && xp > 0)) return false;
  *rest_min = mn > 0 ? -INFINITY : (mp > 0 ? INFINITY : vmin);
  *rest_max = xn > 0 ? -INFINITY : (xp > 0 ? INFINITY : vmax);
  return !isnan(*rest_min) && !isnan(*rest_max);
}

// flags[0]: some live singleton column; flags[1]: some on an equality row;
// flags[2]: some on a row that is not an equality.
__global__ void k_cs_precheck(int32_t n, int32_t m, gpk_plan p, const uint8_t* mask, const int32_t* srow,
                              double tol, uint8_t* needed, int32_t* flags) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || !p.keep_col[j] || !mask[j]) return;
  const int32_t row = srow[j];
  if (row < 0 || row >= m || !p.keep_row[row]) return;
  needed[row] = 1;
  wave_set_flag(&flags[0]);
  const double lo = p.AL[row], hi = p.AU[row];
  if (isfinite(lo) && isfinite(hi) && fabs(hi - lo) <= tol) wave_set_flag(&flags[1]);
  else wave_set_flag(&flags[2]);
}

// rule_column_singletons.cu:455-535 (sequential along the row).
// The row is walked U entries at a time: their columns and values load
// together, then the columns' keep flags and bounds, and the terms are then
// added in CSR order (the sums are those of the entry-by-entry loop).
__global__ __launch_bounds__(kBlock) void k_cs_activity_cache(gpk_csr A, gpk_plan p, const uint8_t* needed,
                                                              gpk_row_cache rc) {
  constexpr int U = 8;
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.rows || !needed[i]) return;
  double fmn = 0.0, fmx = 0.0;
  int32_t mn = 0, mp = 0, xn = 0, xp = 0;
  bool invalid = false;
  if (p.keep_row[i]) {
    const int32_t qb = A.ptr[i], qe = A.ptr[i + 1];
    for (int32_t q0 = qb; q0 < qe; q0 += U) {
      int32_t cc[U];
      double av[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int32_t q = min(q0 + k, qe - 1);
        cc[k] = A.idx[q];
        av[k] = A.val[q];
      }
      bool kept[U];
      double lv[U], uv[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        kept[k] = p.keep_col[cc[k]] != 0;
        lv[k] = p.l[cc[k]];
        uv[k] = p.u[cc[k]];
      }
#pragma unroll
      for (int k = 0; k < U; ++k) {
        if (q0 + k >= qe || !kept[k]) continue;
        const double a = av[k];
        const double tmin = a >= 0.0 ? a * lv[k] : a * uv[k];
        const double tmax = a >= 0.0 ? a * uv[k] : a * lv[k];
        if (isnan(tmin) || isnan(tmax)) {
          invalid = true;
          continue;
        }
        if (isinf(tmin)) {
          if (tmin < 0.0) ++mn; else ++mp;
        } else {
          fmn += tmin;
        }
        if (isinf(tmax)) {
          if (tmax < 0.0) ++xn; else ++xp;
        } else {
          fmx += tmax;
        }
      }
    }
  }
  rc.fmin[i] = fmn;
  rc.fmax[i] = fmx;
  rc.mn[i] = mn;
  rc.mp[i] = mp;
  rc.xn[i] = xn;
  rc.xp[i] = xp;
  rc.invalid[i] = invalid ? 1 : 0;
}

// Shared qualification test of the owner and process kernels.
__device__ bool cs_qualifies(const gpk_plan& p, int32_t j, int32_t row, double a, double lhs, double rhs,
                             const gpk_row_cache& rc, double tol, double zt, int32_t eq_mode, bool* is_eq,
                             bool* above, bool* below, double* rmin, double* rmax) {
  if (!cs_rest_bounds(rc, row, j, a, p.l, p.u, rmin, rmax)) return false;
  *is_eq = isfinite(lhs) && isfinite(rhs) && fabs(lhs - rhs) <= tol;
  if (*is_eq) {
    if (!eq_mode) return false;
    const double x1 = (rhs - *rmin) / a, x2 = (rhs - *rmax) / a;
    const double ilb = fmin(x1, x2), iub = fmax(x1, x2);
    *above = !isfinite(p.u[j]) || iub <= p.u[j] + tol;
    *below = !isfinite(p.l[j]) || ilb >= p.l[j] - tol;
    return *above || *below;
  }
  if (eq_mode) return false;
  *above = cs_free_above(a, lhs, rhs, p.u[j], *rmin, *rmax, tol);
  *below = cs_free_below(a, lhs, rhs, p.l[j], *rmin, *rmax, tol);
  double side;
  return cs_ineq_action(p.c[j], a, lhs, rhs, *above, *below, zt, &side) != kIneqNone;
}

// rule_column_singletons.cu:802-921; flags[0] unbounded.
__global__ void k_cs_owner(gpk_csr A, gpk_plan p, const uint8_t* mask, const int32_t* srow, const double* sval,
                           gpk_row_cache rc, double tol, double zt, int32_t eq_mode, int32_t* owner,
                           int32_t* flags) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t m = A.rows;
  if (j >= A.cols || !p.keep_col[j] || !mask[j]) return;
  const int32_t row = srow[j];
  if (row < 0 || row >= m || !p.keep_row[row]) return;
  const double a = sval[j];
  if (fabs(a) <= zt) return;
  int32_t live = 0;
  for (int32_t q = A.ptr[row]; q < A.ptr[row + 1]; ++q)
    if (p.keep_col[A.idx[q]] && ++live > 1) break;
  if (live <= 1) return;
  const double lhs = p.AL[row], rhs = p.AU[row], cj = p.c[j];
  if (!eq_mode && cs_direct_unbounded(cj, a, lhs, rhs, p.l[j], p.u[j], zt)) {
    wave_set_flag(&flags[0]);
    return;
  }
  bool is_eq, above, below;
  double rmin, rmax;
  if (!cs_qualifies(p, j, row, a, lhs, rhs, rc, tol, zt, eq_mode, &is_eq, &above, &below, &rmin, &rmax)) return;
  atomicMin(&owner[row], j);
}

// rule_column_singletons.cu:923-1061; flags[0] unbounded, [1] EQ eliminations, [2] dual changes.
__global__ void k_cs_process(int32_t n, int32_t m, gpk_plan p, const uint8_t* mask, const int32_t* srow,
                             const double* sval, const int32_t* owner, gpk_row_cache rc, double tol, double zt,
                             int32_t eq_mode, double* AL_new, double* AU_new, uint8_t* row_del,
                             uint8_t* row_changed, uint8_t* col_del, double* side_of, double* contrib,
                             int32_t* flags) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || !p.keep_col[j] || !mask[j]) return;
  const int32_t row = srow[j];
  if (row < 0 || row >= m || !p.keep_row[row] || owner[row] != j) return;
  const double a = sval[j];
  if (fabs(a) <= zt) return;
  const double lhs = AL_new[row], rhs = AU_new[row], cj = p.c[j];
  if (!eq_mode && cs_direct_unbounded(cj, a, lhs, rhs, p.l[j], p.u[j], zt)) {
    wave_set_flag(&flags[0]);
    return;
  }
  bool is_eq, above, below;
  double rmin, rmax;
  if (!cs_qualifies(p, j, row, a, lhs, rhs, rc, tol, zt, eq_mode, &is_eq, &above, &below, &rmin, &rmax)) return;
  if (is_eq) {
    side_of[j] = rhs;
    contrib[j] = cj * rhs / a;
    if (above && below) {
      row_del[row] = 1;
    } else {
      const double bval = above ? p.l[j] : p.u[j];
      const bool keep_lower = above ? (a < 0.0) : (a > 0.0);
      const double shifted = AU_new[row] - a * bval;
      if (keep_lower) {
        AL_new[row] = shifted;
        AU_new[row] = INFINITY;
      } else {
        AL_new[row] = -INFINITY;
        AU_new[row] = shifted;
      }
    }
    col_del[j] = 1;
    wave_set_flag(&flags[1]);
    return;
  }
  double side = 0.0;
  const uint8_t act = cs_ineq_action(cj, a, lhs, rhs, above, below, zt, &side);
  if (act == kIneqElim) {
    if (!isfinite(lhs) || fabs(lhs - side) > tol) row_changed[row] = 1;
    if (!isfinite(rhs) || fabs(rhs - side) > tol) row_changed[row] = 1;
    AL_new[row] = side;
    AU_new[row] = side;
  } else if (act == kIneqLhsToRhs) {
    row_changed[row] = 1;
    AL_new[row] = rhs;
  } else {
    row_changed[row] = 1;
    AU_new[row] = lhs;
  }
  wave_set_flag(&flags[2]);
}

// rule_column_singletons.cu:1063-1091 (AT order, no atomics).
__global__ void k_cs_objective(gpk_csr AT, gpk_plan p, const uint8_t* col_del, const int32_t* owner,
                               const double* sval, double* c_new) {
  const int32_t col = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t n = AT.rows;
  if (col >= n) return;
  double upd = p.c[col];
  if (p.keep_col[col] && !col_del[col]) {
    for (int32_t q = AT.ptr[col]; q < AT.ptr[col + 1]; ++q) {
      const int32_t ow = owner[AT.idx[q]];
      if (ow < 0 || ow >= n || !col_del[ow]) continue;
      upd += -(p.c[ow] * AT.val[q] / sval[ow]);
    }
  }
  c_new[col] = upd;
}

__global__ void k_cs_subcol_counts(gpk_csr A, gpk_plan p, const int32_t* owner, const uint8_t* col_del,
                                   uint8_t* selected, int32_t* support) {
  const int32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= A.rows) return;
  const int32_t col = owner[row];
  const bool sel = col >= 0 && col < A.cols && col_del[col];
  selected[row] = sel ? 1 : 0;
  int32_t cnt = 0;
  if (sel)
    for (int32_t q = A.ptr[row]; q < A.ptr[row + 1]; ++q) {
      const int32_t rc2 = A.idx[q];
      cnt += (rc2 != col && p.keep_col[rc2]) ? 1 : 0;
    }
  support[row] = cnt;
}

// out_i2 (3 per record): {col, row, support start}; out_v6: {pivot, side, l, u, c, row_deleted}.
__global__ void k_cs_subcol_pack(gpk_csr A, gpk_plan p, const int32_t* owner, const uint8_t* row_del,
                                 const double* side_of, const uint8_t* selected, const int32_t* sel_prefix,
                                 const int32_t* sup_prefix, int32_t* out_i2, double* out_v6, int32_t* out_scol,
                                 double* out_sval) {
  const int32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= A.rows || !selected[row]) return;
  const int32_t col = owner[row];
  const int32_t r = sel_prefix[row];
  int32_t o = sup_prefix[row];
  double pivot = 0.0;
  for (int32_t q = A.ptr[row]; q < A.ptr[row + 1]; ++q) {
    const int32_t rc2 = A.idx[q];
    if (rc2 == col) {
      pivot = A.val[q];
    } else if (p.keep_col[rc2]) {
      out_scol[o] = rc2;
      out_sval[o] = A.val[q];
      ++o;
    }
  }
  out_i2[3 * r] = col;
  out_i2[3 * r + 1] = row;
  out_i2[3 * r + 2] = sup_prefix[row];
  out_v6[6 * static_cast<int64_t>(r) + 0] = pivot;
  out_v6[6 * static_cast<int64_t>(r) + 1] = side_of[col];
  out_v6[6 * static_cast<int64_t>(r) + 2] = p.l[col];
  out_v6[6 * static_cast<int64_t>(r) + 3] = p.u[col];
  out_v6[6 * static_cast<int64_t>(r) + 4] = p.c[col];
  out_v6[6 * static_cast<int64_t>(r) + 5] = row_del[row] ? 1.0 : 0.0;
}

__global__ void k_cs_apply(int32_t m, int32_t n, gpk_plan p, const uint8_t* row_del, const uint8_t* col_del,
                           const double* AL_new, const double* AU_new, const double* c_new) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < m) {
    if (row_del[q]) p.keep_row[q] = 0;
    p.AL[q] = AL_new[q];
    p.AU[q] = AU_new[q];
  }
  if (q < n) {
    if (col_del[q]) p.keep_col[q] = 0;
    p.c[q] = c_new[q];
  }
}

// -------------------------------------------------------- dual fix ----
// flags[0]: atomicMax of 1 (infeasible) / 2 (unbounded); flags[1] any fix; flags[2] any infinite fix.
__global__ void k_dual_fix_candidates(gpk_csr AT, gpk_plan p, const uint8_t* cand, double zt, double btol,
                                      int32_t max_deg, uint8_t* fixed, uint8_t* inf_fix, uint8_t* finite_fix,
                                      double* v, double* contrib, int32_t* flags) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= AT.rows) return;
  fixed[j] = 0;
  inf_fix[j] = 0;
  finite_fix[j] = 0;
  v[j] = 0.0;
  contrib[j] = 0.0;
  if (!p.keep_col[j] || (cand != nullptr && !cand[j])) return;
  const double lj = p.l[j], uj = p.u[j];
  if (lj > uj + btol) {
    wave_max(&flags[0], 1);
    return;
  }
  const int32_t b = AT.ptr[j], e = AT.ptr[j + 1];
  if (fabs(p.c[j]) > zt && e - b > max_deg) return;
  bool down = false, up = false;
  for (int32_t q = b; q < e; ++q) {
    const int32_t row = AT.idx[q];
    if (!p.keep_row[row]) continue;
    const double a = AT.val[q];
    if (a > zt) {
      down = down || isfinite(p.AL[row]);
      up = up || isfinite(p.AU[row]);
    } else if (a < -zt) {
      down = down || isfinite(p.AU[row]);
      up = up || isfinite(p.AL[row]);
    }
    if (down && up) break;
  }
  const double cj = p.c[j];
  bool fx = false, infx = false;
  double vj = 0.0;
  if (cj > zt) {
    if (!down) {
      if (isfinite(lj)) { vj = lj; fx = true; }
      else { wave_max(&flags[0], 2); return; }
    }
  } else if (cj < -zt) {
    if (!up) {
      if (isfinite(uj)) { vj = uj; fx = true; }
      else { wave_max(&flags[0], 2); return; }
    }
  } else if (fabs(cj) <= zt) {
    if (!down) {
      fx = true;
      if (isfinite(lj)) vj = lj; else { vj = -INFINITY; infx = true; }
    } else if (!up) {
      fx = true;
      if (isfinite(uj)) vj = uj; else { vj = INFINITY; infx = true; }
    }
  }
  if (fx) {
    fixed[j] = 1;
    inf_fix[j] = infx ? 1 : 0;
    finite_fix[j] = infx ? 0 : 1;
    v[j] = vj;
    contrib[j] = infx ? 0.0 : cj * vj;
    wave_set_flag(&flags[1]);
    if (infx) wave_set_flag(&flags[2]);
  }
}

// Per live row: shift of the finite fixes (CSR order) and removal when the
// row holds an infinitely fixed column; flags[3] row removed.
__global__ void k_dual_fix_rows(gpk_csr A, gpk_plan p, const uint8_t* finite_fix, const uint8_t* inf_fix,
                                const double* v, double* shift, uint8_t* remove, int32_t* flags) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.rows) return;
  double s = 0.0;
  uint8_t rm = 0;
  if (p.keep_row[i]) {
    for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
      const int32_t j = A.idx[q];
      if (finite_fix[j]) s += A.val[q] * v[j];
      if (inf_fix[j]) rm = 1;
    }
  }
  shift[i] = s;
  remove[i] = rm;
  if (rm) wave_set_flag(&flags[3]);
}

__global__ void k_dual_fix_apply(int32_t m, int32_t n, gpk_plan p, const uint8_t* fixed, const double* v,
                                 const double* shift, const uint8_t* remove) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n && fixed[k]) {
    p.l[k] = v[k];
    p.u[k] = v[k];
    p.keep_col[k] = 0;
  }
  if (k < m) {
    p.AL[k] -= shift[k];
    p.AU[k] -= shift[k];
    if (remove[k]) p.keep_row[k] = 0;
  }
}

// ------------------------------------------------ duplicate columns ----
constexpr uint64_t kFnvOffset = 0xcbf29ce484222325ULL;
constexpr uint64_t kFnvPrime = 0x100000001b3ULL;
__device__ inline uint64_t fnv_mix(uint64_t h, uint64_t x) { return (h ^ x) * kFnvPrime; }
__device__ inline uint64_t dbits(double v) { return static_cast<uint64_t>(__double_as_longlong(v)); }

// rule_duplicate_columns.cu:126-172
__global__ void k_dup_col_hash(gpk_csr AT, gpk_plan p, double zt, uint64_t* hash) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= AT.rows) return;
  if (!p.keep_col[j]) {
    hash[j] = static_cast<uint64_t>(j);
    return;
  }
  int32_t live = 0;
  double piv = 0.0;
  bool set = false;
  for (int32_t q = AT.ptr[j]; q < AT.ptr[j + 1]; ++q) {
    if (p.keep_row[AT.idx[q]] && fabs(AT.val[q]) > zt) {
      ++live;
      if (!set) { piv = AT.val[q]; set = true; }
    }
  }
  if (!set || live == 0) {
    hash[j] = fnv_mix(kFnvOffset, static_cast<uint64_t>(j));
    return;
  }
  uint64_t h = fnv_mix(kFnvOffset, static_cast<uint64_t>(live));
  for (int32_t q = AT.ptr[j]; q < AT.ptr[j + 1]; ++q) {
    if (p.keep_row[AT.idx[q]] && fabs(AT.val[q]) > zt) {
      h = fnv_mix(h, static_cast<uint64_t>(static_cast<int64_t>(AT.idx[q])));
      h = fnv_mix(h, dbits(AT.val[q] / piv));
    }
  }
  hash[j] = h;
}

__device__ inline int32_t dc_next_live(const gpk_csr& AT, const gpk_plan& p, double zt, int32_t q, int32_t stop) {
  while (q < stop) {
    if (p.keep_row[AT.idx[q]] && fabs(AT.val[q]) > zt) return q;
    ++q;
  }
  return stop;
}

// rule_duplicate_columns.cu:181-229
__device__ bool dc_ratio(const gpk_csr& AT, const gpk_plan& p, int32_t j, int32_t k, double zt, double ctol,
                         double* out) {
  int32_t pj = AT.ptr[j], sj = AT.ptr[j + 1], pk = AT.ptr[k], sk = AT.ptr[k + 1];
  double r = 0.0;
  bool set = false;
  for (;;) {
    pj = dc_next_live(AT, p, zt, pj, sj);
    pk = dc_next_live(AT, p, zt, pk, sk);
    if (pj >= sj || pk >= sk) break;
    if (AT.idx[pj] != AT.idx[pk]) return false;
    const double aj = AT.val[pj], ak = AT.val[pk];
    if (!set) {
      if (fabs(aj) <= zt) return false;
      r = ak / aj;
      if (fabs(r) <= zt || !isfinite(r)) return false;
      set = true;
    }
    if (fabs(ak - r * aj) > ctol) return false;
    ++pj;
    ++pk;
  }
  pj = dc_next_live(AT, p, zt, pj, sj);
  pk = dc_next_live(AT, p, zt, pk, sk);
  if (pj < sj || pk < sk || !set) return false;
  *out = r;
  return true;
}

__device__ inline bool ratio_violates(double amax, double r) {
  const double ar = fabs(r);
  return amax < 1.0 || ar > amax || ar * amax < 1.0;
}

// rule_duplicate_columns.cu:284-344: number of small buckets holding a
// cost-parallel pair whose ratio breaks the guard.
__global__ void k_dup_col_guard(gpk_csr AT, gpk_plan p, const uint64_t* sh, const int32_t* sc, double zt,
                                double ctol, double otol, double max_ratio, int32_t max_group, int32_t* count) {
  const int32_t n = AT.rows;
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint64_t hs = sh[s];
  if (s > 0 && sh[s - 1] == hs) return;
  int32_t e = s;
  while (e + 1 < n && sh[e + 1] == hs) ++e;
  const int32_t size = e - s + 1;
  if (size < 2 || size > max_group) return;
  for (int32_t a = s; a < e; ++a) {
    const int32_t j = sc[a];
    if (!p.keep_col[j]) continue;
    for (int32_t b = a + 1; b <= e; ++b) {
      const int32_t k = sc[b];
      if (!p.keep_col[k]) continue;
      double r;
      if (!dc_ratio(AT, p, j, k, zt, ctol, &r)) continue;
      if (fabs(p.c[k] - r * p.c[j]) > otol) continue;
      if (ratio_violates(max_ratio, r)) {
        wave_add1(count);
        return;
      }
    }
  }
}

// rule_duplicate_columns.cu:346-497, one thread per bucket.  flags[0]
// unbounded, [1] merged, [2] fixed.  rec5[5*k]: ratio, deleted_l, deleted_u, kept_l, kept_u.
__global__ void k_dup_col_groups(gpk_csr AT, gpk_plan p, const uint64_t* sh, const int32_t* sc, double zt,
                                 double ctol, double otol, double max_ratio, const int32_t* guard_count,
                                 int32_t guard_min, uint8_t* col_del, uint8_t* fixed, double* fixed_val,
                                 int32_t* merge_to, double* rec5, int32_t* flags) {
  const int32_t n = AT.rows;
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint64_t hs = sh[s];
  if (s > 0 && sh[s - 1] == hs) return;
  int32_t e = s;
  while (e + 1 < n && sh[e + 1] == hs) ++e;
  if (e <= s) return;
  double amax = max_ratio;
  if (guard_count != nullptr && *guard_count < guard_min) amax = INFINITY;
  for (int32_t a = s; a < e; ++a) {
    const int32_t j = sc[a];
    if (!p.keep_col[j] || col_del[j] || fixed[j]) continue;
    for (int32_t b = a + 1; b <= e; ++b) {
      const int32_t k = sc[b];
      if (!p.keep_col[k] || col_del[k] || fixed[k]) continue;
      double r;
      if (!dc_ratio(AT, p, j, k, zt, ctol, &r)) continue;
      if (isfinite(amax) && ratio_violates(amax, r)) continue;
      const double gap = p.c[k] - r * p.c[j];
      const double tl = p.l[j], tu = p.u[j], sl = p.l[k], su = p.u[k];
      if (fabs(gap) <= otol) {
        merge_to[k] = j;
        rec5[5 * static_cast<int64_t>(k) + 0] = r;
        rec5[5 * static_cast<int64_t>(k) + 1] = sl;
        rec5[5 * static_cast<int64_t>(k) + 2] = su;
        rec5[5 * static_cast<int64_t>(k) + 3] = tl;
        rec5[5 * static_cast<int64_t>(k) + 4] = tu;
        double nl = -INFINITY, nu = INFINITY;
        if (r > 0.0) {
          if (isfinite(tl) && isfinite(sl)) nl = tl + r * sl;
          if (isfinite(tu) && isfinite(su)) nu = tu + r * su;
        } else {
          if (isfinite(tl) && isfinite(su)) nl = tl + r * su;
          if (isfinite(tu) && isfinite(sl)) nu = tu + r * sl;
        }
        p.l[j] = nl;
        p.u[j] = nu;
        col_del[k] = 1;
        wave_set_flag(&flags[1]);
        continue;
      }
      bool k_lo = false, k_hi = false, j_lo = false, j_hi = false;
      if (gap > otol) {
        if (r > 0.0) { k_lo = !isfinite(tu); j_hi = !isfinite(sl); }
        else { k_lo = !isfinite(tl); j_lo = !isfinite(sl); }
      } else {
        if (r > 0.0) { k_hi = !isfinite(tl); j_lo = !isfinite(su); }
        else { k_hi = !isfinite(tu); j_hi = !isfinite(su); }
      }
      if (k_lo) {
        if (!isfinite(sl)) { wave_set_flag(&flags[0]); return; }
        fixed[k] = 1; fixed_val[k] = sl; wave_set_flag(&flags[2]); continue;
      } else if (k_hi) {
        if (!isfinite(su)) { wave_set_flag(&flags[0]); return; }
        fixed[k] = 1; fixed_val[k] = su; wave_set_flag(&flags[2]); continue;
      }
      if (j_lo) {
        if (!isfinite(tl)) { wave_set_flag(&flags[0]); return; }
        fixed[j] = 1; fixed_val[j] = tl; wave_set_flag(&flags[2]); break;
      } else if (j_hi) {
        if (!isfinite(tu)) { wave_set_flag(&flags[0]); return; }
        fixed[j] = 1; fixed_val[j] = tu; wave_set_flag(&flags[2]); break;
      }
    }
  }
}

__global__ void k_dup_col_apply(int32_t n, gpk_plan p, const uint8_t* col_del, const uint8_t* fixed,
                                const double* fixed_val, double* contrib) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  contrib[j] = 0.0;
  if (!(col_del[j] || fixed[j])) return;
  p.keep_col[j] = 0;
  if (fixed[j]) {
    p.l[j] = fixed_val[j];
    p.u[j] = fixed_val[j];
    contrib[j] = p.c[j] * fixed_val[j];
  }
}

// ------------------------------------------------ doubleton (batch) ----
constexpr double kDblMaxRatio = 1.0e3;

__device__ inline bool dbl_integral_ratio(double num, double den, double tol) {
  if (fabs(den) <= tol) return false;
  const double r = fabs(num / den);
  if (!isfinite(r)) return false;
  return fabs(r - nearbyint(r)) <= tol;
}

__device__ inline bool dbl_pivot_ok(double keep_val, double elim_val) {
  if (fabs(elim_val) <= 0.0) return false;
  const double r = fabs(keep_val / elim_val);
  return r <= kDblMaxRatio && r >= 1.0 / kDblMaxRatio;
}

__device__ inline int32_t dbl_next_live(const gpk_csr& AT, const uint8_t* keep_row, double zt, int32_t col,
                                        int32_t* pos) {
  const int32_t stop = AT.ptr[col + 1];
  while (*pos < stop) {
    const int32_t r = AT.idx[*pos];
    if (keep_row[r] && fabs(AT.val[*pos]) > zt) return r;
    ++(*pos);
  }
  return INT_MAX;
}

__device__ int32_t dbl_fill_in_live(const gpk_csr& AT, const uint8_t* keep_row, double zt, int32_t keep,
                                    int32_t elim) {
  int32_t fill = -1;
  int32_t jj = AT.ptr[keep], ks = AT.ptr[keep + 1], kk = AT.ptr[elim], es = AT.ptr[elim + 1];
  if (ks <= jj || es <= kk) return fill;
  while (jj < ks && kk < es) {
    int32_t jl = jj, kl = kk;
    const int32_t kr = dbl_next_live(AT, keep_row, zt, keep, &jl);
    const int32_t er = dbl_next_live(AT, keep_row, zt, elim, &kl);
    if (kr == INT_MAX || er == INT_MAX) break;
    if (kr == er) { jj = jl + 1; kk = kl + 1; }
    else if (er < kr) { kk = kl + 1; ++fill; }
    else jj = jl + 1;
  }
  while (kk < es) {
    int32_t kl = kk;
    const int32_t er = dbl_next_live(AT, keep_row, zt, elim, &kl);
    if (er == INT_MAX) break;
    ++fill;
    kk = kl + 1;
  }
  return fill;
}

__device__ int32_t dbl_fill_in_static(const gpk_csr& AT, int32_t keep, int32_t elim) {
  int32_t fill = -1;
  int32_t jj = AT.ptr[keep], ks = AT.ptr[keep + 1], kk = AT.ptr[elim], es = AT.ptr[elim + 1];
  if (ks <= jj || es <= kk) return fill;
  while (jj < ks && kk < es) {
    const int32_t kr = AT.idx[jj], er = AT.idx[kk];
    if (kr == er) { ++jj; ++kk; }
    else if (er < kr) { ++kk; ++fill; }
    else ++jj;
  }
  return fill + (es - kk);
}

__global__ void k_dbl_live_nnz(gpk_csr AT, gpk_plan p, double zt, int32_t* nnz) {
  const int32_t col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= AT.rows) return;
  int32_t cnt = 0;
  if (p.keep_col[col])
    for (int32_t q = AT.ptr[col]; q < AT.ptr[col + 1]; ++q)
      cnt += (p.keep_row[AT.idx[q]] && fabs(AT.val[q]) > zt) ? 1 : 0;
  nnz[col] = cnt;
}

// Live two-entry equality rows -> (elim, keep) choice and fill-in gate
// (rule_doubleton_equations.cu:164-236, 746-874).  counts[0]: acceptable rows.
__global__ void k_dbl_candidates(gpk_csr A, gpk_csr AT, gpk_plan p, gpk_dbl d, double zt, double tol, double rtol,
                                 int32_t max_fill, int32_t live_fill, int32_t* counts) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.rows) return;
  d.acceptable[i] = 0;
  d.active[i] = 0;
  d.selected[i] = 0;
  if (!p.keep_row[i]) return;
  const double lhs = p.AL[i], rhs = p.AU[i];
  if (!isfinite(lhs) || !isfinite(rhs) || fabs(lhs - rhs) > tol) return;
  int32_t live = 0, c1 = -1, c2 = -1;
  double v1 = 0.0, v2 = 0.0;
  for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
    const int32_t col = A.idx[q];
    const double a = A.val[q];
    if (p.keep_col[col] && fabs(a) > zt) {
      ++live;
      if (live == 1) { c1 = col; v1 = a; }
      else if (live == 2) { c2 = col; v2 = a; }
      else return;
    }
  }
  if (live != 2) return;
  const int32_t* nnz = d.live_nnz;
  const bool i12 = dbl_integral_ratio(v1, v2, rtol), i21 = dbl_integral_ratio(v2, v1, rtol);
  bool elim_first;
  if (nnz[c1] == 1 && nnz[c2] != 1) elim_first = true;
  else if (nnz[c1] != 1 && nnz[c2] == 1) elim_first = false;
  else if (i12 && !i21) elim_first = false;
  else if (i21 && !i12) elim_first = true;
  else if (nnz[c1] < nnz[c2]) elim_first = true;
  else elim_first = false;
  const int32_t ec = elim_first ? c1 : c2, kc = elim_first ? c2 : c1;
  const double ev = elim_first ? v1 : v2, kv = elim_first ? v2 : v1;
  if (!dbl_pivot_ok(kv, ev)) return;
  d.cec[i] = ec;
  d.ckc[i] = kc;
  d.cev[i] = ev;
  d.ckv[i] = kv;
  if (nnz[ec] > nnz[kc] + max_fill + 1) return;
  const int32_t fill = live_fill ? dbl_fill_in_live(AT, p.keep_row, zt, kc, ec) : dbl_fill_in_static(AT, kc, ec);
  if (fill > max_fill) return;
  d.acceptable[i] = 1;
  d.active[i] = 1;
  wave_add1(&counts[0]);
}

// A claim of column ownership in `round` by row i: larger for a later round,
// and within a round larger for a lower row, so an atomic max keeps the
// lowest claiming row of the latest round (a minimum-row claim, as the
// reference's per-round atomicMin after a reset).
__device__ inline unsigned long long dbl_stamp(int32_t round, int32_t i) {
  return (static_cast<unsigned long long>(round) << 32) | (0xffffffffull - static_cast<uint32_t>(i));
}

__global__ void k_dbl_claim(int32_t m, gpk_dbl d, int32_t round) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !d.active[i]) return;
  const int32_t k = d.ckc[i], e = d.cec[i];
  if (d.blocked[k] || d.blocked[e]) return;
  const unsigned long long v = dbl_stamp(round, i);
  atomicMax(reinterpret_cast<unsigned long long*>(&d.owner[k]), v);
  atomicMax(reinterpret_cast<unsigned long long*>(&d.owner[e]), v);
}

// counts[1]: selected rows.  A row wins when it owns both of its columns in
// this round; owners are unique, so winners never share a column.
__global__ void k_dbl_select(int32_t m, gpk_dbl d, int32_t round, int32_t* counts) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !d.active[i]) return;
  const int32_t k = d.ckc[i], e = d.cec[i];
  const uint64_t v = dbl_stamp(round, i);
  if (!d.blocked[k] && !d.blocked[e] && d.owner[k] == v && d.owner[e] == v) {
    d.selected[i] = 1;
    d.active[i] = 0;
    d.blocked[k] = 1;
    d.blocked[e] = 1;
    wave_add1(&counts[1]);
  }
}

// rule_doubleton_equations.cu:916-1004 plus tape capture.  flags[0]
// infeasible, [2] substitutions, [3] fixed keep columns.
// rec_i3: {elim, keep, row}; rec_v12: DoubletonEquation values.
__global__ void k_dbl_apply(int32_t k, const int32_t* sel, gpk_plan p, gpk_dbl d, gpk_dbl_cols cols, double tol,
                            uint8_t* subst, uint8_t* fixedp, double* osub, double* ofix, int32_t* rec_i3,
                            double* rec_v12, int32_t* flags) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k) return;
  const int32_t row = sel[t], kc = d.ckc[row], ec = d.cec[row];
  const double kv = d.ckv[row], ev = d.cev[row], rhs = p.AU[row];
  const double ol_e = p.l[ec], ou_e = p.u[ec], oc_e = p.c[ec], ol_k = p.l[kc], ou_k = p.u[kc], oc_k = p.c[kc];
  subst[t] = 0;
  fixedp[t] = 0;
  osub[t] = 0.0;
  ofix[t] = 0.0;
  double* rv = rec_v12 + 12 * static_cast<int64_t>(t);
  rec_i3[3 * t + 0] = ec;
  rec_i3[3 * t + 1] = kc;
  rec_i3[3 * t + 2] = row;
  rv[0] = ev; rv[1] = kv; rv[2] = rhs; rv[3] = ol_e; rv[4] = ou_e; rv[5] = ol_k; rv[6] = ou_k;
  rv[7] = 0.0; rv[8] = 0.0; rv[9] = oc_e; rv[10] = 0.0; rv[11] = 0.0;
  const double al = -kv / ev, be = rhs / ev;
  double ml = -INFINITY, mu = INFINITY;
  if (isfinite(ol_e)) { const double b = (ol_e - be) / al; if (al > 0.0) ml = b; else mu = b; }
  if (isfinite(ou_e)) { const double b = (ou_e - be) / al; if (al > 0.0) mu = b; else ml = b; }
  if (ml > mu) { const double tmp = ml; ml = mu; mu = tmp; }
  const double nkl = isfinite(ml) ? fmax(ol_k, ml) : ol_k;
  const double nku = isfinite(mu) ? fmin(ou_k, mu) : ou_k;
  if (nkl > nku + tol) {
    wave_set_flag(&flags[0]);
    return;
  }
  const bool keep_fixed = isfinite(nkl) && isfinite(nku) && nku <= nkl + tol;
  if (keep_fixed) {
    const double fat = 0.5 * (nkl + nku);
    p.l[kc] = fat;
    p.u[kc] = fat;
    p.keep_col[kc] = 0;
    fixedp[t] = 1;
    ofix[t] = oc_k * fat;
    cols.is_fixed_keep[kc] = 1;
    cols.fixed_val_of[kc] = fat;
    wave_add1(&flags[3]);
  } else {
    p.l[kc] = nkl;
    p.u[kc] = nku;
    p.c[kc] = oc_k + oc_e * al;
    p.keep_row[row] = 0;
    p.keep_col[ec] = 0;
    subst[t] = 1;
    osub[t] = oc_e * be;
    rv[7] = nkl;
    rv[8] = nku;
    cols.is_elim[ec] = 1;
    cols.alpha_of[ec] = al;
    cols.beta_of[ec] = be;
    cols.row_of[ec] = row;
    cols.keep_of[ec] = kc;
    wave_add1(&flags[2]);
  }
}

__device__ inline bool dbl_is_delta(const gpk_dbl_cols& cols, int32_t i, int32_t j, double a, double zt) {
  return cols.is_elim[j] && cols.row_of[j] != i && fabs(a) > zt;
}

__global__ void k_dbl_delta_count(gpk_csr A, gpk_plan p, gpk_dbl_cols cols, double zt, int32_t* counts) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.rows) return;
  int32_t cnt = 0;
  if (p.keep_row[i])
    for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) cnt += dbl_is_delta(cols, i, A.idx[q], A.val[q], zt) ? 1 : 0;
  counts[i] = cnt;
}

// Row shift of one batch: fixed keep columns first, then substituted columns,
// each in CSR order along the row; applied to every row.
__global__ void k_dbl_shift(gpk_csr A, gpk_plan p, gpk_dbl_cols cols, double zt, int32_t include_subst) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.rows) return;
  double s = 0.0;
  if (p.keep_row[i]) {
    for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
      const int32_t j = A.idx[q];
      if (cols.is_fixed_keep[j] && fabs(A.val[q]) > zt) s += A.val[q] * cols.fixed_val_of[j];
    }
    if (include_subst)
      for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
        const int32_t j = A.idx[q];
        if (dbl_is_delta(cols, i, j, A.val[q], zt)) s += A.val[q] * cols.beta_of[j];
      }
  }
  p.AL[i] -= s;
  p.AU[i] -= s;
}

// Delta entries (keep column, alpha * a) of each live row, sorted by column.
__global__ void k_dbl_delta_fill(gpk_csr A, gpk_plan p, gpk_dbl_cols cols, double zt, const int32_t* dptr,
                                 int32_t* dcol, double* dval) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.rows || !p.keep_row[i]) return;
  const int32_t b = dptr[i];
  int32_t o = b;
  for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
    const int32_t j = A.idx[q];
    if (!dbl_is_delta(cols, i, j, A.val[q], zt)) continue;
    const int32_t kc = cols.keep_of[j];
    const double v = cols.alpha_of[j] * A.val[q];
    int32_t w = o - 1;
    while (w >= b && dcol[w] > kc) {
      dcol[w + 1] = dcol[w];
      dval[w + 1] = dval[w];
      --w;
    }
    dcol[w + 1] = kc;
    dval[w + 1] = v;
    ++o;
  }
}

// Sorted merge of a row with its deltas (rule_doubleton_equations.cu:1129-1218);
// count mode when out_idx is null.
__global__ void k_dbl_merge(gpk_csr A, const int32_t* dptr, const int32_t* dcol, const double* dval, double zt,
                            const int32_t* out_ptr, int32_t* counts, int32_t* out_idx, double* out_val) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.rows) return;
  int32_t q = A.ptr[i];
  const int32_t qe = A.ptr[i + 1];
  int32_t r = dptr[i];
  const int32_t re = dptr[i + 1];
  int32_t w = out_idx ? out_ptr[i] : 0;
  int32_t cnt = 0;
  while (q < qe || r < re) {
    const int32_t sc = q < qe ? A.idx[q] : INT_MAX;
    const int32_t dc = r < re ? dcol[r] : INT_MAX;
    const int32_t nc = sc < dc ? sc : dc;
    double merged = 0.0;
    while (q < qe && A.idx[q] == nc) { merged += A.val[q]; ++q; }
    while (r < re && dcol[r] == nc) { merged += dval[r]; ++r; }
    if (fabs(merged) > zt) {
      if (out_idx) {
        out_idx[w] = nc;
        out_val[w] = merged;
        ++w;
      }
      ++cnt;
    }
  }
  if (!out_idx) counts[i] = cnt;
}


// Single-doubleton rebuild of A (rule_doubleton_equations.cu:1094-1127,
// 1779-1858): every row but the target gets keep += alpha * a_elim; the keep
// entry is inserted in column order or dropped when it becomes tiny.  Count
// mode when out_idx is null.
__global__ void k_dbl_scalar_rows(gpk_csr A, int32_t target, int32_t ec, int32_t kc, double al, double zt,
                                  const int32_t* out_ptr, int32_t* counts, int32_t* out_idx, double* out_val) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= A.rows) return;
  const int32_t b = A.ptr[r], e = A.ptr[r + 1];
  int32_t w = out_idx ? out_ptr[r] : 0;
  int32_t cnt = 0;
  auto emit = [&](int32_t col, double v) {
    if (out_idx) {
      out_idx[w] = col;
      out_val[w] = v;
      ++w;
    }
    ++cnt;
  };
  if (r == target) {
    for (int32_t q = b; q < e; ++q) emit(A.idx[q], A.val[q]);
  } else {
    bool ep = false, kp = false;
    double are = 0.0, okv = 0.0;
    for (int32_t q = b; q < e; ++q) {
      if (A.idx[q] == ec) { ep = true; are = A.val[q]; }
      else if (A.idx[q] == kc) { kp = true; okv = A.val[q]; }
    }
    const double nk = okv + al * are;
    const bool insert = ep && !kp && fabs(nk) > zt;
    bool inserted = !insert;
    for (int32_t q = b; q < e; ++q) {
      const int32_t col = A.idx[q];
      if (insert && !inserted && col > kc) { emit(kc, nk); inserted = true; }
      if (col == kc) {
        if (fabs(nk) > zt) emit(kc, nk);
      } else {
        emit(col, A.val[q]);
      }
    }
    if (insert && !inserted) emit(kc, nk);
  }
  if (!out_idx) counts[r] = cnt;
}


// Row shift of the single-doubleton path: every row t holding the eliminated
// column in the phase-entry AT gets AL/AU -= a_te * beta.
__global__ void k_dbl_scalar_shift(gpk_csr AT, int32_t ec, double beta, double* AL, double* AU) {
  const int32_t q = AT.ptr[ec] + blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= AT.ptr[ec + 1]) return;
  const int32_t t = AT.idx[q];
  const double sh = AT.val[q] * beta;
  AL[t] -= sh;
  AU[t] -= sh;
}

// ------------------------------------------------ redundant bounds ----
__device__ bool rb_rest(const gpk_csr& A, const gpk_plan& p, int32_t i, int32_t j, bool want_min, double* out) {
  double acc = 0.0;
  for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
    const int32_t col = A.idx[q];
    if (col == j) continue;
    const double aq = A.val[q];
    const double t = want_min ? (aq >= 0.0 ? aq * p.l[col] : aq * p.u[col])
                              : (aq >= 0.0 ? aq * p.u[col] : aq * p.l[col]);
    if (!isfinite(t)) return false;
    acc += t;
  }
  *out = acc;
  return true;
}

__global__ void k_rb_candidates(gpk_csr A, gpk_csr AT, gpk_plan p, const double* c_lp, int32_t one_sided, double zt,
                                double tol, uint8_t* dlo, uint8_t* dhi) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= AT.rows) return;
  dlo[j] = 0;
  dhi[j] = 0;
  const double lj = p.l[j], uj = p.u[j];
  const bool box = isfinite(lj) && isfinite(uj);
  const bool allow_box = box && (c_lp == nullptr || fabs(c_lp[j]) <= zt);
  const bool eligible = allow_box || (one_sided && !box);
  if (isfinite(uj) && eligible) {
    for (int32_t q = AT.ptr[j]; q < AT.ptr[j + 1]; ++q) {
      const int32_t i = AT.idx[q];
      const double a = AT.val[q];
      if (fabs(a) <= zt) continue;
      double iu = INFINITY, r;
      if (a > zt && isfinite(p.AU[i])) { if (rb_rest(A, p, i, j, true, &r)) iu = (p.AU[i] - r) / a; }
      else if (a < -zt && isfinite(p.AL[i])) { if (rb_rest(A, p, i, j, false, &r)) iu = (p.AL[i] - r) / a; }
      if (isfinite(iu) && iu <= uj + tol) { dhi[j] = 1; break; }
    }
  }
  if (isfinite(lj) && eligible) {
    for (int32_t q = AT.ptr[j]; q < AT.ptr[j + 1]; ++q) {
      const int32_t i = AT.idx[q];
      const double a = AT.val[q];
      if (fabs(a) <= zt) continue;
      double il = -INFINITY, r;
      if (a > zt && isfinite(p.AL[i])) { if (rb_rest(A, p, i, j, false, &r)) il = (p.AL[i] - r) / a; }
      else if (a < -zt && isfinite(p.AU[i])) { if (rb_rest(A, p, i, j, true, &r)) il = (p.AU[i] - r) / a; }
      if (isfinite(il) && il >= lj - tol) { dlo[j] = 1; break; }
    }
  }
}

__device__ inline uint64_t rb_key(const gpk_csr& AT, int32_t j) {
  return (static_cast<uint64_t>(static_cast<uint32_t>(AT.ptr[j + 1] - AT.ptr[j])) << 32) |
         static_cast<uint64_t>(static_cast<uint32_t>(j));
}

__global__ void k_rb_row_key(gpk_csr A, gpk_csr AT, const uint8_t* dlo, const uint8_t* dhi, uint64_t* rkey) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.rows) return;
  uint64_t best = ~0ull;
  for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
    const int32_t j = A.idx[q];
    if (dlo[j] || dhi[j]) {
      const uint64_t k = rb_key(AT, j);
      if (k < best) best = k;
    }
  }
  rkey[i] = best;
}

__global__ void k_rb_select(gpk_csr AT, const uint8_t* dlo, const uint8_t* dhi, const uint64_t* rkey, uint8_t* sl,
                            uint8_t* su, int32_t* flags) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= AT.rows) return;
  sl[j] = 0;
  su[j] = 0;
  if (!dlo[j] && !dhi[j]) return;
  const uint64_t key = rb_key(AT, j);
  for (int32_t q = AT.ptr[j]; q < AT.ptr[j + 1]; ++q)
    if (rkey[AT.idx[q]] != key) return;
  sl[j] = dlo[j];
  su[j] = dhi[j];
  wave_set_flag(&flags[0]);
}

__global__ void k_rb_apply(int32_t n, gpk_plan p, const uint8_t* sl, const uint8_t* su) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if (sl[j]) p.l[j] = -INFINITY;
  if (su[j]) p.u[j] = INFINITY;
}

}  // namespace

#define GPK_LAUNCH(kernel, n, ...)                                                      \
  do {                                                                                  \
    if ((n) > 0) hipLaunchKernelGGL(kernel, dim3(grid_for(n)), dim3(kBlock), 0, stream, __VA_ARGS__); \
    return hipGetLastError();                                                           \
  } while (0)

extern "C" {

hipError_t gpk_fixed_vars_candidates(int32_t n, gpk_plan plan, double bound_tol, uint8_t* fixed, double* v,
                                     double* contrib, int32_t* count, hipStream_t stream) {
  GPK_LAUNCH(k_fixed_vars_candidates, n, n, plan, bound_tol, fixed, v, contrib, count);
}
hipError_t gpk_empty_cols(int32_t n, gpk_plan plan, const uint8_t* empty_mask, double bound_tol, double zero_tol,
                          uint8_t* fixed, double* v, double* contrib, int32_t* flags, hipStream_t stream) {
  GPK_LAUNCH(k_empty_cols, n, n, plan, empty_mask, bound_tol, zero_tol, fixed, v, contrib, flags);
}
hipError_t gpk_fix_cols_apply(int32_t n, gpk_plan plan, const uint8_t* fixed, const double* v, hipStream_t stream) {
  GPK_LAUNCH(k_fix_cols_apply, n, n, plan, fixed, v);
}
hipError_t gpk_cs_precheck(int32_t n, int32_t m, gpk_plan plan, const uint8_t* mask, const int32_t* srow,
                           double bound_tol, uint8_t* needed, int32_t* flags, hipStream_t stream) {
  GPK_LAUNCH(k_cs_precheck, n, n, m, plan, mask, srow, bound_tol, needed, flags);
}
hipError_t gpk_cs_activity_cache(gpk_csr A, gpk_plan plan, const uint8_t* needed, gpk_row_cache cache,
                                 hipStream_t stream) {
  GPK_LAUNCH(k_cs_activity_cache, A.rows, A, plan, needed, cache);
}
hipError_t gpk_cs_owner(gpk_csr A, gpk_plan plan, const uint8_t* mask, const int32_t* srow, const double* sval,
                        gpk_row_cache cache, double tol, double zero_tol, int32_t eq_mode, int32_t* owner,
                        int32_t* flags, hipStream_t stream) {
  GPK_LAUNCH(k_cs_owner, A.cols, A, plan, mask, srow, sval, cache, tol, zero_tol, eq_mode, owner, flags);
}
hipError_t gpk_cs_process(int32_t n, int32_t m, gpk_plan plan, const uint8_t* mask, const int32_t* srow,
                          const double* sval, const int32_t* owner, gpk_row_cache cache, double tol,
                          double zero_tol, int32_t eq_mode, double* AL_new, double* AU_new, uint8_t* row_del,
                          uint8_t* row_changed, uint8_t* col_del, double* side, double* contrib, int32_t* flags,
                          hipStream_t stream) {
  GPK_LAUNCH(k_cs_process, n, n, m, plan, mask, srow, sval, owner, cache, tol, zero_tol, eq_mode, AL_new, AU_new,
             row_del, row_changed, col_del, side, contrib, flags);
}
hipError_t gpk_cs_objective(gpk_csr AT, gpk_plan plan, const uint8_t* col_del, const int32_t* owner,
                            const double* sval, double* c_new, hipStream_t stream) {
  GPK_LAUNCH(k_cs_objective, AT.rows, AT, plan, col_del, owner, sval, c_new);
}
hipError_t gpk_cs_subcol_counts(gpk_csr A, gpk_plan plan, const int32_t* owner, const uint8_t* col_del,
                                uint8_t* selected, int32_t* support_count, hipStream_t stream) {
  GPK_LAUNCH(k_cs_subcol_counts, A.rows, A, plan, owner, col_del, selected, support_count);
}
hipError_t gpk_cs_subcol_pack(gpk_csr A, gpk_plan plan, const int32_t* owner, const uint8_t* row_del,
                              const double* side, const uint8_t* selected, const int32_t* sel_prefix,
                              const int32_t* sup_prefix, int32_t* out_i2, double* out_v6, int32_t* out_scol,
                              double* out_sval, hipStream_t stream) {
  GPK_LAUNCH(k_cs_subcol_pack, A.rows, A, plan, owner, row_del, side, selected, sel_prefix, sup_prefix, out_i2,
             out_v6, out_scol, out_sval);
}
hipError_t gpk_cs_apply(int32_t m, int32_t n, gpk_plan plan, const uint8_t* row_del, const uint8_t* col_del,
                        const double* AL_new, const double* AU_new, const double* c_new, hipStream_t stream) {
  const int32_t mn = m > n ? m : n;
  GPK_LAUNCH(k_cs_apply, mn, m, n, plan, row_del, col_del, AL_new, AU_new, c_new);
}
hipError_t gpk_dual_fix_candidates(gpk_csr AT, gpk_plan plan, const uint8_t* candidate, double zero_tol,
                                   double bound_tol, int32_t max_degree, uint8_t* fixed, uint8_t* inf_fix,
                                   uint8_t* finite_fix, double* v, double* contrib, int32_t* flags,
                                   hipStream_t stream) {
  GPK_LAUNCH(k_dual_fix_candidates, AT.rows, AT, plan, candidate, zero_tol, bound_tol, max_degree, fixed, inf_fix,
             finite_fix, v, contrib, flags);
}
hipError_t gpk_dual_fix_rows(gpk_csr A, gpk_plan plan, const uint8_t* finite_fix, const uint8_t* inf_fix,
                             const double* v, double* shift, uint8_t* row_remove, int32_t* flags,
                             hipStream_t stream) {
  GPK_LAUNCH(k_dual_fix_rows, A.rows, A, plan, finite_fix, inf_fix, v, shift, row_remove, flags);
}
hipError_t gpk_dual_fix_apply(int32_t m, int32_t n, gpk_plan plan, const uint8_t* fixed, const double* v,
                              const double* shift, const uint8_t* row_remove, hipStream_t stream) {
  const int32_t mn = m > n ? m : n;
  GPK_LAUNCH(k_dual_fix_apply, mn, m, n, plan, fixed, v, shift, row_remove);
}
hipError_t gpk_dup_col_hash(gpk_csr AT, gpk_plan plan, double zero_tol, uint64_t* hash, hipStream_t stream) {
  GPK_LAUNCH(k_dup_col_hash, AT.rows, AT, plan, zero_tol, hash);
}
hipError_t gpk_dup_col_guard(gpk_csr AT, gpk_plan plan, const uint64_t* sorted_hash, const int32_t* sorted_cols,
                             double zero_tol, double coeff_tol, double obj_tol, double max_ratio, int32_t max_group,
                             int32_t* count, hipStream_t stream) {
  GPK_LAUNCH(k_dup_col_guard, AT.rows, AT, plan, sorted_hash, sorted_cols, zero_tol, coeff_tol, obj_tol, max_ratio,
             max_group, count);
}
hipError_t gpk_dup_col_groups(gpk_csr AT, gpk_plan plan, const uint64_t* sorted_hash, const int32_t* sorted_cols,
                              double zero_tol, double coeff_tol, double obj_tol, double max_ratio,
                              const int32_t* guard_count, int32_t guard_min_groups, uint8_t* col_del,
                              uint8_t* fixed, double* fixed_val, int32_t* merge_to, double* rec5, int32_t* flags,
                              hipStream_t stream) {
  GPK_LAUNCH(k_dup_col_groups, AT.rows, AT, plan, sorted_hash, sorted_cols, zero_tol, coeff_tol, obj_tol, max_ratio,
             guard_count, guard_min_groups, col_del, fixed, fixed_val, merge_to, rec5, flags);
}
hipError_t gpk_dup_col_apply(int32_t n, gpk_plan plan, const uint8_t* col_del, const uint8_t* fixed,
                             const double* fixed_val, double* contrib, hipStream_t stream) {
  GPK_LAUNCH(k_dup_col_apply, n, n, plan, col_del, fixed, fixed_val, contrib);
}
hipError_t gpk_dbl_live_nnz(gpk_csr AT, gpk_plan plan, double zero_tol, int32_t* live_nnz, hipStream_t stream) {
  GPK_LAUNCH(k_dbl_live_nnz, AT.rows, AT, plan, zero_tol, live_nnz);
}
hipError_t gpk_dbl_candidates(gpk_csr A, gpk_csr AT, gpk_plan plan, gpk_dbl d, double zero_tol, double tol,
                              double ratio_tol, int32_t max_fill, int32_t live_fill, int32_t* counts,
                              hipStream_t stream) {
  GPK_LAUNCH(k_dbl_candidates, A.rows, A, AT, plan, d, zero_tol, tol, ratio_tol, max_fill, live_fill, counts);
}
hipError_t gpk_dbl_claim(int32_t m, gpk_dbl d, int32_t round, hipStream_t stream) {
  GPK_LAUNCH(k_dbl_claim, m, m, d, round);
}
hipError_t gpk_dbl_select(int32_t m, gpk_dbl d, int32_t round, int32_t* counts, hipStream_t stream) {
  GPK_LAUNCH(k_dbl_select, m, m, d, round, counts);
}
hipError_t gpk_dbl_apply(int32_t k, const int32_t* sel, gpk_plan plan, gpk_dbl d, gpk_dbl_cols cols, double tol,
                         uint8_t* subst, uint8_t* fixedp, double* osub, double* ofix, int32_t* rec_i3,
                         double* rec_v12, int32_t* flags, hipStream_t stream) {
  GPK_LAUNCH(k_dbl_apply, k, k, sel, plan, d, cols, tol, subst, fixedp, osub, ofix, rec_i3, rec_v12, flags);
}
hipError_t gpk_dbl_delta_count(gpk_csr A, gpk_plan plan, gpk_dbl_cols cols, double zero_tol, int32_t* counts,
                               hipStream_t stream) {
  GPK_LAUNCH(k_dbl_delta_count, A.rows, A, plan, cols, zero_tol, counts);
}
hipError_t gpk_dbl_shift(gpk_csr A, gpk_plan plan, gpk_dbl_cols cols, double zero_tol, int32_t include_subst,
                         hipStream_t stream) {
  GPK_LAUNCH(k_dbl_shift, A.rows, A, plan, cols, zero_tol, include_subst);
}
hipError_t gpk_dbl_delta_fill(gpk_csr A, gpk_plan plan, gpk_dbl_cols cols, double zero_tol, const int32_t* delta_ptr,
                              int32_t* dcol, double* dval, hipStream_t stream) {
  GPK_LAUNCH(k_dbl_delta_fill, A.rows, A, plan, cols, zero_tol, delta_ptr, dcol, dval);
}
hipError_t gpk_dbl_merge(gpk_csr A, const int32_t* delta_ptr, const int32_t* dcol, const double* dval,
                         double zero_tol, const int32_t* out_ptr, int32_t* counts, int32_t* out_idx, double* out_val,
                         hipStream_t stream) {
  GPK_LAUNCH(k_dbl_merge, A.rows, A, delta_ptr, dcol, dval, zero_tol, out_ptr, counts, out_idx, out_val);
}
hipError_t gpk_dbl_scalar_rows(gpk_csr A, int32_t target, int32_t ec, int32_t kc, double alpha, double zero_tol,
                               const int32_t* out_ptr, int32_t* counts, int32_t* out_idx, double* out_val,
                               hipStream_t stream) {
  GPK_LAUNCH(k_dbl_scalar_rows, A.rows, A, target, ec, kc, alpha, zero_tol, out_ptr, counts, out_idx, out_val);
}
hipError_t gpk_dbl_scalar_shift(gpk_csr AT, int32_t ec, int32_t ec_len, double beta, double* AL, double* AU,
                                hipStream_t stream) {
  GPK_LAUNCH(k_dbl_scalar_shift, ec_len, AT, ec, beta, AL, AU);
}
hipError_t gpk_rb_candidates(gpk_csr A, gpk_csr AT, gpk_plan plan, const double* c_lp, int32_t one_sided,
                             double zero_tol, double tol, uint8_t* dlo, uint8_t* dhi, hipStream_t stream) {
  GPK_LAUNCH(k_rb_candidates, AT.rows, A, AT, plan, c_lp, one_sided, zero_tol, tol, dlo, dhi);
}
hipError_t gpk_rb_select(gpk_csr A, gpk_csr AT, const uint8_t* dlo, const uint8_t* dhi, uint64_t* row_key,
                         uint8_t* sel_lo, uint8_t* sel_hi, int32_t* flags, hipStream_t stream) {
  if (A.rows > 0) hipLaunchKernelGGL(k_rb_row_key, dim3(grid_for(A.rows)), dim3(kBlock), 0, stream, A, AT, dlo, dhi,
                                     row_key);
  GPK_LAUNCH(k_rb_select, AT.rows, AT, dlo, dhi, row_key, sel_lo, sel_hi, flags);
}
hipError_t gpk_rb_apply(int32_t n, gpk_plan plan, const uint8_t* sel_lo, const uint8_t* sel_hi, hipStream_t stream) {
  GPK_LAUNCH(k_rb_apply, n, n, plan, sel_lo, sel_hi);
}

}  // extern "C"
