// CPU oracle: activity primitives and the reduction rules -- TEST INFRASTRUCTURE.
// Each function cites the reference source it restates (paths relative to
// PolyU-IOR/GPU-Presolver/cpp/src/presolve/).
#include "oracle_core.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

namespace {

inline bool fin(double x) { return std::isfinite(x); }

// Semantics of the reference's CAS loops atomic_max_double/atomic_min_double
// (rules/rule_singleton_rows.cu:18-50): the slot is replaced unless it already
// compares >= (resp. <=) the value.
inline void cas_max(double& slot, double v) {
  if (!(slot >= v)) slot = v;
}
inline void cas_min(double& slot, double v) {
  if (!(slot <= v)) slot = v;
}

inline uint64_t dbits(double v) {
  uint64_t b;
  std::memcpy(&b, &v, sizeof b);
  return b;
}

constexpr uint64_t FNV_OFFSET = 0xcbf29ce484222325ULL;
constexpr uint64_t FNV_PRIME = 0x100000001b3ULL;
inline uint64_t fnv_mix(uint64_t h, uint64_t x) { return (h ^ x) * FNV_PRIME; }

// Sequential sum in ascending index order: the canonical order for the
// objective deltas the reference sums with cub::DeviceReduce::Sum.
double ordered_sum(const std::vector<double>& v) {
  double s = 0.0;
  for (double x : v) s += x;
  return s;
}

}  // namespace

void default_params(Params* p) {
  std::memset(p, 0, sizeof(*p));
  p->max_iters = 10;
  p->max_time = INF;
  p->verbose = 0;
  p->record_postsolve_tape = 1;
  p->record_postsolve_tape_cpu = 0;
  p->feasibility_tol = 1.0e-6;
  p->bound_tol = 1.0e-6;
  p->zero_tol = 1.0e-10;
  p->implied_variable_bounds_max_rounds = 64;
  p->implied_variable_bounds_max_bound_only_rounds = 64;
  p->implied_variable_bounds_preserve_free_zero_cost = 1;
  p->doubleton_equations_single_batch_per_iter = 0;
  p->doubleton_equations_max_fill_in_proxy = 96;
  p->doubleton_equations_scan = 1;
  p->doubleton_equations_min_selected_per_batch = 256;
  p->doubleton_equations_min_selected_ratio = 0.005;
  p->doubleton_equations_max_batch_rounds = 0;
  p->doubleton_equations_max_time = 0.0;
  p->enable_infeasible_fixed_variables = 1;
  p->enable_empty_rows = 1;
  p->enable_singleton_rows = 1;
  p->enable_infeasible_redundant_rows = 1;
  p->enable_implied_variable_bounds = 1;
  p->enable_duplicate_rows = 1;
  p->enable_empty_cols = 1;
  p->enable_column_singletons_eq = 1;
  p->enable_column_singletons_dual_infer = 1;
  p->enable_doubleton_equations = 1;
  p->enable_dual_fix = 1;
  p->dual_fix_max_nonzero_cost_col_degree = 128;
  p->enable_duplicate_columns = 1;
  p->duplicate_columns_max_abs_ratio = 2.0;
  p->duplicate_columns_scale_guard_max_group_size = 4;
  p->duplicate_columns_scale_guard_min_groups = 128;
  p->enable_structure_specific_rules = 1;
  p->enable_zero_cost_redundant_box_bounds = 0;
  p->enable_redundant_bounds = 0;
  p->use_tiered_scheduler = 1;
  p->enable_tiered_bootstrap = 1;
  p->implied_variable_bounds_bound_only_nnz_round_budget = 8000000;
  // The structure-specific rules (presolve_structs.hpp:75-115): the oracle
  // does not run them; the fields only keep the struct at its defaults.
  p->enable_bounded_two_row_projection = 1;
  p->bounded_two_row_projection_max_row_nnz = 8;
  p->bounded_two_row_projection_safety_tol = 1.0e-8;
  p->bounded_two_row_projection_min_candidates = 128;
  p->bounded_two_row_projection_min_candidate_ratio = 0.01;
  p->enable_orphan_mccormick_projection = 1;
  p->orphan_mccormick_projection_min_candidates = 128;
  p->orphan_mccormick_projection_min_candidate_ratio = 0.01;
  p->enable_linf_components = 1;
  p->linf_probe_pairs = 256;
  p->linf_min_pairs = 4096;
  p->linf_max_tail_cols = 8;
  p->enable_structural_l1_substitution = 1;
  p->enable_antipodal_components = 1;
  p->antipodal_probe_pairs = 256;
  p->antipodal_min_pairs = 4096;
  p->enable_covering_cost_dominance = 1;
  p->covering_cost_dominance_min_fixed_cols = 4096;
  p->covering_cost_dominance_min_candidate_ratio = 0.20;
}

Csr copy_csr(const CsrV& a) {
  Csr out;
  out.rows = a.rows;
  out.cols = a.cols;
  out.ptr.assign(a.ptr, a.ptr + a.rows + 1);
  out.idx.assign(a.idx, a.idx + a.nnz);
  out.val.assign(a.val, a.val + a.nnz);
  return out;
}

#ifdef _OPENMP
// All-core builds only (the CPU baseline's OpenMP leg): the same transpose
// and compaction with the work split over threads; outputs identical to the
// serial versions below.  Rows are split into one contiguous block per
// thread; each thread counts its block's entries per column, a per-column
// exclusive sum over the threads gives every thread its own cursors, and
// the blocks are scattered in row order, so every column keeps ascending
// rows.
Csr transpose_threads(const CsrV& a) {
  const int T = omp_get_max_threads();
  const int64_t n = a.cols;
  Csr t;
  t.rows = a.cols;
  t.cols = a.rows;
  t.ptr.assign(static_cast<size_t>(n) + 1, 0);
  t.idx.resize(a.nnz);
  t.val.resize(a.nnz);
  std::vector<int32_t> cnt(static_cast<size_t>(T) * static_cast<size_t>(n), 0);
  auto block = [&](int k) { return static_cast<int32_t>(static_cast<int64_t>(a.rows) * k / T); };
#pragma omp parallel num_threads(T)
  {
    const int k = omp_get_thread_num();
    int32_t* c = cnt.data() + static_cast<size_t>(k) * static_cast<size_t>(n);
    for (int32_t p = a.ptr[block(k)]; p < a.ptr[block(k + 1)]; ++p) ++c[a.idx[p]];
  }
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < n; ++j) {
    int32_t run = 0;
    for (int k = 0; k < T; ++k) {
      int32_t& c = cnt[static_cast<size_t>(k) * static_cast<size_t>(n) + static_cast<size_t>(j)];
      const int32_t v = c;
      c = run;
      run += v;
    }
    t.ptr[j + 1] = run;
  }
  for (int64_t j = 0; j < n; ++j) t.ptr[j + 1] += t.ptr[j];
#pragma omp parallel num_threads(T)
  {
    const int k = omp_get_thread_num();
    int32_t* c = cnt.data() + static_cast<size_t>(k) * static_cast<size_t>(n);
    for (int32_t i = block(k); i < block(k + 1); ++i)
      for (int32_t p = a.ptr[i]; p < a.ptr[i + 1]; ++p) {
        const int32_t j = a.idx[p];
        const int32_t q = t.ptr[j] + c[j]++;
        t.idx[q] = i;
        t.val[q] = a.val[p];
      }
  }
  return t;
}

// Kept entries per kept row counted in parallel, a serial prefix over the
// rows, then the kept entries written in parallel.
Csr compact_threads(const CsrV& a, const uint8_t* keep_row, const uint8_t* keep_col) {
  Csr out;
  std::vector<int32_t> col_new(a.cols, -1);
  int32_t n_new = 0;
  for (int32_t j = 0; j < a.cols; ++j) col_new[j] = keep_col[j] ? n_new++ : -1;
  out.cols = n_new;
  std::vector<int32_t> row_new(a.rows, -1);
  int32_t m_new = 0;
  for (int32_t i = 0; i < a.rows; ++i) row_new[i] = keep_row[i] ? m_new++ : -1;
  out.rows = m_new;
  out.ptr.assign(static_cast<size_t>(m_new) + 1, 0);
  if (n_new > 0) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < a.rows; ++i) {
      if (row_new[i] < 0) continue;
      int32_t k = 0;
      for (int32_t p = a.ptr[i]; p < a.ptr[i + 1]; ++p) k += col_new[a.idx[p]] >= 0;
      out.ptr[row_new[i] + 1] = k;
    }
  }
  for (int32_t r = 0; r < m_new; ++r) out.ptr[r + 1] += out.ptr[r];
  out.idx.resize(out.ptr[m_new]);
  out.val.resize(out.ptr[m_new]);
  if (n_new > 0) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < a.rows; ++i) {
      if (row_new[i] < 0) continue;
      int32_t w = out.ptr[row_new[i]];
      for (int32_t p = a.ptr[i]; p < a.ptr[i + 1]; ++p) {
        const int32_t jn = col_new[a.idx[p]];
        if (jn >= 0) {
          out.idx[w] = jn;
          out.val[w] = a.val[p];
          ++w;
        }
      }
    }
  }
  return out;
}
#endif

// Transpose with rows of the result in ascending row-of-A order, the output
// of cusparseCsr2cscEx2 ALG1 used by the reference (gpu_presolve.cu:1804-1879).
Csr transpose(const CsrV& a) {
#ifdef _OPENMP
  if (omp_get_max_threads() > 1 && a.nnz > 65536) return transpose_threads(a);
#endif
  Csr t;
  t.rows = a.cols;
  t.cols = a.rows;
  t.ptr.assign(static_cast<size_t>(a.cols) + 1, 0);
  t.idx.resize(a.nnz);
  t.val.resize(a.nnz);
  for (int32_t p = 0; p < a.nnz; ++p) t.ptr[a.idx[p] + 1]++;
  for (int32_t j = 0; j < a.cols; ++j) t.ptr[j + 1] += t.ptr[j];
  std::vector<int32_t> cursor(t.ptr.begin(), t.ptr.end() - 1);
  for (int32_t i = 0; i < a.rows; ++i) {
    for (int32_t p = a.ptr[i]; p < a.ptr[i + 1]; ++p) {
      const int32_t q = cursor[a.idx[p]]++;
      t.idx[q] = i;
      t.val[q] = a.val[p];
    }
  }
  return t;
}

// Stable row/column compaction (gpu_presolve.cu:297-531, 1881-1978).
Csr compact(const CsrV& a, const uint8_t* keep_row, const uint8_t* keep_col) {
#ifdef _OPENMP
  if (omp_get_max_threads() > 1 && a.nnz > 65536) return compact_threads(a, keep_row, keep_col);
#endif
  Csr out;
  std::vector<int32_t> col_new(a.cols, -1);
  int32_t n_new = 0;
  for (int32_t j = 0; j < a.cols; ++j) col_new[j] = keep_col[j] ? n_new++ : -1;
  out.cols = n_new;
  out.ptr.assign(1, 0);
  int32_t m_new = 0;
  for (int32_t i = 0; i < a.rows; ++i) {
    if (!keep_row[i]) continue;
    ++m_new;
    // The reference writes an empty matrix (all-zero rowPtr) when either
    // dimension becomes zero (gpu_presolve.cu:1890-1894).
    if (n_new > 0) {
      for (int32_t p = a.ptr[i]; p < a.ptr[i + 1]; ++p) {
        const int32_t jn = col_new[a.idx[p]];
        if (jn >= 0) {
          out.idx.push_back(jn);
          out.val.push_back(a.val[p]);
        }
      }
    }
    out.ptr.push_back(static_cast<int32_t>(out.idx.size()));
  }
  out.rows = m_new;
  return out;
}

// Row sort of a CSR in place by column, values following, as the reference
// does to the caller's AT before the doubleton batch and duplicate-column
// rules (rule_doubleton_equations.cu:95-141, rule_duplicate_columns.cu:575-595).
void sort_rows_in_place(const CsrV& a) {
  auto* idx = const_cast<int32_t*>(a.idx);
  auto* val = const_cast<double*>(a.val);
  std::vector<std::pair<int32_t, double>> tmp;
  for (int32_t r = 0; r < a.rows; ++r) {
    const int32_t b = a.ptr[r], e = a.ptr[r + 1];
    if (std::is_sorted(idx + b, idx + e)) continue;
    tmp.clear();
    for (int32_t p = b; p < e; ++p) tmp.emplace_back(idx[p], val[p]);
    std::stable_sort(tmp.begin(), tmp.end(),
                     [](const auto& x, const auto& y) { return x.first < y.first; });
    for (int32_t p = b; p < e; ++p) {
      idx[p] = tmp[p - b].first;
      val[p] = tmp[p - b].second;
    }
  }
}

// gpu_presolve_kernels.cu:18-66
void row_activity(const CsrV& A, const double* l, const double* u, double ztol, double* min_fin,
                  double* max_fin, int32_t* neg_inf, int32_t* pos_inf) {
#pragma omp parallel for schedule(static) if (A.nnz > 65536)
  for (int32_t i = 0; i < A.rows; ++i) {
    double mn = 0.0, mx = 0.0;
    int32_t ni = 0, pi = 0;
    for (int32_t p = A.ptr[i]; p < A.ptr[i + 1]; ++p) {
      const double a = A.val[p];
      if (std::fabs(a) <= ztol) continue;
      const int32_t j = A.idx[p];
      const double tmin = a >= 0.0 ? a * l[j] : a * u[j];
      const double tmax = a >= 0.0 ? a * u[j] : a * l[j];
      if (fin(tmin)) mn += tmin; else if (tmin < 0.0) ++ni;
      if (fin(tmax)) mx += tmax; else if (tmax > 0.0) ++pi;
    }
    min_fin[i] = mn;
    max_fin[i] = mx;
    neg_inf[i] = ni;
    pos_inf[i] = pi;
  }
}

// gpu_presolve_kernels.cu:68-82
void col_max_abs(const CsrV& AT, double* out) {
  for (int32_t j = 0; j < AT.rows; ++j) {
    double v = 0.0;
    for (int32_t p = AT.ptr[j]; p < AT.ptr[j + 1]; ++p) v = std::fmax(v, std::fabs(AT.val[p]));
    out[j] = v;
  }
}

// fixed_col_tape.cu:107-267: FixedCol records in ascending column order,
// objective delta accumulated in long double.
double fixed_col_tape(Tape* tape, const uint8_t* fixed_mask, const uint8_t* excluded,
                      const double* fixed_val, const uint8_t* keep_row, const double* c,
                      const CsrV& AT, int32_t n) {
  long double obj = 0.0L;
  for (int32_t j = 0; j < n; ++j) {
    if (!fixed_mask[j] || (excluded && excluded[j])) continue;
    const double cj = c ? c[j] : 0.0;
    obj += static_cast<long double>(cj) * static_cast<long double>(fixed_val[j]);
    if (!tape) continue;
    std::vector<int32_t> idx{j};
    std::vector<double> v{fixed_val[j], cj};
    for (int32_t p = AT.ptr[j]; p < AT.ptr[j + 1]; ++p) {
      if (keep_row[AT.idx[p]]) {
        idx.push_back(AT.idx[p]);
        v.push_back(AT.val[p]);
      }
    }
    tape->append(FixedCol, idx, v, DualMinimal);
  }
  return static_cast<double>(obj);
}

// ---------------------------------------------------------------------------
// rules/rule_empty_rows.cu:143-240
void rule_empty_rows(Plan& p, const LpV& lp, const Stats& s, const Params& prm) {
  if (p.has_infeasible || p.has_unbounded) return;
  const int32_t m = lp.A.rows;
  if (m == 0) return;
  const double tol = prm.feasibility_tol;
  bool infeasible = false, removable = false;
  for (int32_t i = 0; i < m; ++i) {
    if (p.keep_row[i] && s.row_nnz[i] == 0) {
      if (p.AL[i] <= tol && p.AU[i] >= -tol) removable = true; else infeasible = true;
    }
  }
  if (infeasible) { p.has_infeasible = true; return; }
  if (removable) {
    for (int32_t i = 0; i < m; ++i)
      if (p.keep_row[i] && s.row_nnz[i] == 0 && p.AL[i] <= tol && p.AU[i] >= -tol) p.keep_row[i] = 0;
    p.has_row_action = true;
    p.has_change = true;
  }
}

// rules/rule_singleton_rows.cu:54-248
void rule_singleton_rows(Plan& p, const LpV& lp, const Stats& s, const Params& prm) {
  if (p.has_infeasible || p.has_unbounded) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0) return;
  std::vector<double> cl(p.l, p.l + n), cu(p.u, p.u + n);
  std::vector<uint8_t> remove(m, 0);
  for (int32_t i = 0; i < m; ++i) {
    if (!p.keep_row[i] || !s.singleton_row_mask[i]) continue;
    const int32_t col = s.singleton_row_col[i];
    const double a = s.singleton_row_val[i];
    if (col < 0 || col >= n || !(std::fabs(a) > prm.zero_tol)) continue;
    const double lo = p.AL[i], hi = p.AU[i];
    const double il = a > 0.0 ? lo / a : hi / a;
    const double iu = a > 0.0 ? hi / a : lo / a;
    cas_max(cl[col], il);
    cas_min(cu[col], iu);
    remove[i] = 1;
  }
  bool infeasible = false, bounds_changed = false, rows_removed = false;
  for (int32_t j = 0; j < n; ++j) {
    if (cl[j] > cu[j] + prm.bound_tol) infeasible = true;
    if (cl[j] != p.l[j] || cu[j] != p.u[j]) bounds_changed = true;
  }
  for (int32_t i = 0; i < m; ++i) if (remove[i]) rows_removed = true;
  if (infeasible) { p.has_infeasible = true; return; }
  if (!(rows_removed || bounds_changed)) return;
  if (prm.record_postsolve_tape && prm.record_postsolve_tape_cpu && rows_removed) {
    for (int32_t i = 0; i < m; ++i)
      if (remove[i])
        p.tape.append(DeletedRow, {i, s.singleton_row_col[i]},
                      {p.AL[i], p.AU[i], s.singleton_row_val[i]}, DualMinimal);
  }
  for (int32_t i = 0; i < m; ++i) if (remove[i]) p.keep_row[i] = 0;
  for (int32_t j = 0; j < n; ++j) { p.l[j] = cl[j]; p.u[j] = cu[j]; }
  p.has_row_action = true;
  p.has_change = true;
}

// rules/rule_infeasible_redundant_rows.cu:20-243
void rule_infeasible_redundant_rows(Plan& p, const LpV& lp, const Stats& s, const Params& prm) {
  if (p.has_infeasible || p.has_unbounded) return;
  const int32_t m = lp.A.rows;
  if (m == 0) return;
  const double tol = prm.bound_tol;
  std::vector<uint8_t> full(m, 0), dlo(m, 0), dhi(m, 0);
  bool infeasible = false, any = false;
  for (int32_t i = 0; i < m; ++i) {
    if (!p.keep_row[i] || s.row_nnz[i] <= 1) continue;
    const double lo = p.AL[i], hi = p.AU[i];
    const bool lf = fin(lo), hf = fin(hi);
    if (lf && hf && std::fabs(hi - lo) <= tol) continue;
    double rmin = 0.0, rmax = 0.0;
    for (int32_t q = lp.A.ptr[i]; q < lp.A.ptr[i + 1]; ++q) {
      const int32_t j = lp.A.idx[q];
      const double a = lp.A.val[q];
      if (a >= 0.0) { rmin += a * p.l[j]; rmax += a * p.u[j]; }
      else { rmin += a * p.u[j]; rmax += a * p.l[j]; }
    }
    if ((lf && rmax < lo - tol) || (hf && rmin > hi + tol)) { infeasible = true; continue; }
    const bool lo_imp = !lf || rmin >= lo - tol;
    const bool hi_imp = !hf || rmax <= hi + tol;
    if (lo_imp && hi_imp) { full[i] = 1; any = true; }
    else if (lf && lo_imp) { dlo[i] = 1; any = true; }
    else if (hf && hi_imp) { dhi[i] = 1; any = true; }
  }
  if (infeasible) { p.has_infeasible = true; return; }
  if (!any) return;
  if (prm.record_postsolve_tape && prm.record_postsolve_tape_cpu) {
    for (int32_t i = 0; i < m; ++i)
      if (full[i]) p.tape.append(DeletedRow, {i}, {p.AL[i], p.AU[i]}, DualMinimal);
    for (int32_t i = 0; i < m; ++i) {
      if (dlo[i]) p.tape.append(LhsChange, {i}, {p.AL[i], p.AU[i], -INF, p.AU[i]}, DualMinimal);
      if (dhi[i]) p.tape.append(RhsChange, {i}, {p.AL[i], p.AU[i], p.AL[i], INF}, DualMinimal);
    }
  }
  for (int32_t i = 0; i < m; ++i) {
    if (full[i]) p.keep_row[i] = 0;
    if (dlo[i]) p.AL[i] = -INF;
    if (dhi[i]) p.AU[i] = INF;
  }
  p.has_row_action = true;
  p.has_change = true;
}

// ---------------------------------------------------------------------------
// rules/rule_implied_variable_bounds.cu:301-337 (residual activity helpers)
namespace {
inline void term_interval(double a, double lj, double uj, double* tmin, double* tmax) {
  if (a >= 0.0) { *tmin = a * lj; *tmax = a * uj; }
  else { *tmin = a * uj; *tmax = a * lj; }
}
inline double resid_min(double fin_sum, int32_t ninf, double tmin) {
  if (fin(tmin)) return ninf == 0 ? fin_sum - tmin : -INF;
  if (tmin < 0.0) return ninf > 1 ? -INF : fin_sum;
  return fin_sum;
}
inline double resid_max(double fin_sum, int32_t pinf, double tmax) {
  if (fin(tmax)) return pinf == 0 ? fin_sum - tmax : INF;
  if (tmax > 0.0) return pinf > 1 ? INF : fin_sum;
  return fin_sum;
}
// Implied (l, u) of column j from row i for coefficient a
// (rule_implied_variable_bounds.cu:369-393).
inline void implied_from_row(double a, double lo, double hi, double rest_min, double rest_max,
                             double* il, double* iu) {
  *il = -INF;
  *iu = INF;
  if (a > 0.0) {
    if (fin(lo) && fin(rest_max)) *il = (lo - rest_max) / a;
    if (fin(hi) && fin(rest_min)) *iu = (hi - rest_min) / a;
  } else {
    if (fin(hi) && fin(rest_min)) *il = (hi - rest_min) / a;
    if (fin(lo) && fin(rest_max)) *iu = (lo - rest_max) / a;
  }
}
}  // namespace

// rules/rule_implied_variable_bounds.cu:339-1164
void rule_implied_variable_bounds(Plan& p, const LpV& lp, const Stats& s, const Params& prm) {
  if (p.has_infeasible || p.has_unbounded) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0 || n == 0) return;
  const double ztol = prm.zero_tol, ftol = prm.feasibility_tol;
  std::vector<double> cand_l(p.l, p.l + n), cand_u(p.u, p.u + n);
  std::vector<double> fl(p.l, p.l + n), fu(p.u, p.u + n);
  std::vector<double> rmin(m), rmax(m);
  std::vector<int32_t> rni(m), rpi(m);
  row_activity(lp.A, p.l, p.u, ztol, rmin.data(), rmax.data(), rni.data(), rpi.data());
  for (int32_t i = 0; i < m; ++i) {
    if (!p.keep_row[i] || s.row_nnz[i] <= 1) continue;
    for (int32_t q = lp.A.ptr[i]; q < lp.A.ptr[i + 1]; ++q) {
      const int32_t j = lp.A.idx[q];
      const double a = lp.A.val[q];
      if (std::fabs(a) <= ztol) continue;
      const double ol = p.l[j], ou = p.u[j];
      double tmin, tmax;
      term_interval(a, ol, ou, &tmin, &tmax);
      const double rsmin = resid_min(rmin[i], rni[i], tmin);
      const double rsmax = resid_max(rmax[i], rpi[i], tmax);
      double il, iu;
      implied_from_row(a, p.AL[i], p.AU[i], rsmin, rsmax, &il, &iu);
      double nl = ol, nu = ou;
      if (fin(il)) nl = std::fmax(nl, il);
      if (fin(iu)) nu = std::fmin(nu, iu);
      if (nl > ol) cas_max(cand_l[j], nl);
      if (nu < ou) cas_min(cand_u[j], nu);
    }
  }
  std::vector<double> vmax(n);
  col_max_abs(lp.AT, vmax.data());
  const double* ref_l = lp.l ? lp.l : p.l;
  const double* ref_u = lp.u ? lp.u : p.u;
  std::vector<uint8_t> lch(n, 0), uch(n, 0), fixed(n, 0);
  std::vector<double> fixed_at(n, 0.0);
  bool flag_infeasible = false, flag_tight = false, flag_fixed = false;
  // _kernel_finalize_implied_variable_bounds (rule_implied_variable_bounds.cu:513-649)
  for (int32_t j = 0; j < n; ++j) {
    const double lj = p.l[j], uj = p.u[j];
    double cl = cand_l[j], cu = cand_u[j];
    const double vm = vmax[j];
    double nl = lj, nu = uj;
    uint8_t lc = 0, uc = 0, fx = 0;
    double fat = 0.0;
    double el = lj, eu = uj;
    if (fin(cl)) el = std::fmax(el, cl);
    if (fin(cu)) eu = std::fmin(eu, cu);
    if (el > eu + ftol) { flag_infeasible = true; continue; }
    const bool was_free = !fin(ref_l[j]) && !fin(ref_u[j]);
    const bool box = fin(cl) && fin(cu);
    const bool zero_cost = p.c != nullptr && std::fabs(p.c[j]) <= ztol;
    const bool nearly_fixed = box && (cu <= cl || (cu - cl) * vm <= ftol);
    if (prm.implied_variable_bounds_preserve_free_zero_cost && was_free && box && zero_cost &&
        !nearly_fixed) {
      cl = lj;
      cu = uj;
    }
    if (fin(cl) && cl > lj && fin(uj)) {
      if (cl >= uj + ftol) { flag_infeasible = true; continue; }
      if (cl >= uj || (uj - cl) * vm <= ftol) { fx = 1; fat = uj; nl = uj; nu = uj; }
    }
    if (!fx && fin(cl) && cl > lj) {
      const bool tighten = !fin(lj) || ((cl - lj > ftol * 1.0e4) && (cl - lj > 1.0e-2 * std::fabs(lj)));
      if (tighten) {
        if (cl != std::nearbyint(cl)) cl -= 0.5 * ftol * std::fabs(cl);
        nl = cl;
        lc = 1;
      }
    }
    if (!fx && fin(cu) && cu < uj && fin(nl)) {
      if (cu <= nl - ftol) { flag_infeasible = true; continue; }
      if (cu <= nl || (cu - nl) * vm <= ftol) { fx = 1; fat = nl; nu = nl; }
    }
    if (!fx && fin(cu) && cu < uj) {
      const bool tighten = !fin(uj) || ((uj - cu > ftol * 1.0e4) && (uj - cu > 1.0e-2 * std::fabs(uj)));
      if (tighten) {
        if (cu != std::nearbyint(cu)) cu += 0.5 * ftol * std::fabs(cu);
        nu = cu;
        uc = 1;
      }
    }
    if (fx) { lc = 0; uc = 0; flag_fixed = true; }
    if (lc || uc) flag_tight = true;
    fl[j] = nl; fu[j] = nu; lch[j] = lc; uch[j] = uc; fixed[j] = fx; fixed_at[j] = fat;
  }
  if (flag_infeasible) { p.has_infeasible = true; return; }
  if (!(flag_tight || flag_fixed)) return;

  if (prm.record_postsolve_tape && flag_tight) {
    // _kernel_capture_implied_variable_bounds_support_rows_by_col (:668-750):
    // the smallest live row whose implied bound reproduces the new bound.
    const double support_tol = std::fmax(10.0 * prm.bound_tol, 1.0e-10);
    const double tol = prm.bound_tol;
    for (int32_t j = 0; j < n; ++j) {
      int32_t sl = INT_MAX, su = INT_MAX;
      const bool need_l = lch[j] != 0, need_u = uch[j] != 0;
      if (need_l || need_u) {
        const double ol = p.l[j], ou = p.u[j];
        for (int32_t q = lp.AT.ptr[j]; q < lp.AT.ptr[j + 1]; ++q) {
          const int32_t i = lp.AT.idx[q];
          if (!p.keep_row[i] || s.row_nnz[i] <= 1) continue;
          const double a = lp.AT.val[q];
          if (std::fabs(a) <= ztol) continue;
          double tmin, tmax;
          term_interval(a, ol, ou, &tmin, &tmax);
          const double rsmin = resid_min(rmin[i], rni[i], tmin);
          const double rsmax = resid_max(rmax[i], rpi[i], tmax);
          double il, iu;
          implied_from_row(a, p.AL[i], p.AU[i], rsmin, rsmax, &il, &iu);
          if (need_l && fl[j] > ol + tol && fin(il) && std::fabs(il - fl[j]) <= support_tol)
            sl = std::min(sl, i);
          if (need_u && fu[j] < ou - tol && fin(iu) && std::fabs(iu - fu[j]) <= support_tol)
            su = std::min(su, i);
        }
      }
      // _kernel_pack_primal_bound_tape (:771-804)
      double cur_l = p.l[j];
      const double cur_u = p.u[j];
      if (lch[j]) {
        const double tl = fl[j];
        if (sl != INT_MAX) p.tape.append(BoundChangeTheRow, {j, sl}, {cur_l, cur_u, tl, cur_u}, DualMinimal);
        else p.tape.append(BoundChangeNoRow, {j}, {cur_l, cur_u, tl, cur_u}, DualMinimal);
        cur_l = tl;
      }
      if (uch[j]) {
        if (su != INT_MAX) p.tape.append(BoundChangeTheRow, {j, su}, {cur_l, cur_u, cur_l, fu[j]}, DualMinimal);
        else p.tape.append(BoundChangeNoRow, {j}, {cur_l, cur_u, cur_l, fu[j]}, DualMinimal);
      }
    }
  }
  for (int32_t j = 0; j < n; ++j) {
    p.l[j] = fl[j];
    p.u[j] = fu[j];
    if (fixed[j]) p.keep_col[j] = 0;
  }
  if (flag_fixed) {
    // Row shift in canonical per-row ascending-column order (:806-837).
    std::vector<double> shift(m, 0.0);
    for (int32_t i = 0; i < m; ++i) {
      if (!p.keep_row[i]) continue;
      for (int32_t q = lp.A.ptr[i]; q < lp.A.ptr[i + 1]; ++q) {
        const int32_t j = lp.A.idx[q];
        if (fixed[j]) shift[i] += lp.A.val[q] * fixed_at[j];
      }
    }
    for (int32_t i = 0; i < m; ++i) { p.AL[i] -= shift[i]; p.AU[i] -= shift[i]; }
    bool empty_infeasible = false;
    for (int32_t i = 0; i < m; ++i) {
      if (!p.keep_row[i]) continue;
      int32_t cnt = 0;
      for (int32_t q = lp.A.ptr[i]; q < lp.A.ptr[i + 1]; ++q) if (p.keep_col[lp.A.idx[q]]) ++cnt;
      if (cnt == 0) {
        if (p.AL[i] <= ftol && p.AU[i] >= -ftol) p.keep_row[i] = 0;
        else empty_infeasible = true;
      }
    }
    if (empty_infeasible) { p.has_infeasible = true; return; }
  }
  const double obj = flag_fixed
      ? fixed_col_tape(prm.record_postsolve_tape ? &p.tape : nullptr, fixed.data(), nullptr,
                       fixed_at.data(), p.keep_row, p.c, lp.AT, n)
      : 0.0;
  p.obj_delta += obj;
  p.has_row_action = p.has_row_action || flag_fixed;
  p.has_col_action = p.has_col_action || flag_fixed;
  p.has_change = true;
}

// _kernel_duplicate_row_hashes (rules/rule_duplicate_rows.cu:78-114):
// sequential FNV over (len, col, a / a_first); when nnz > 32 m (:484-486)
// the reference switches to _kernel_duplicate_row_hashes_warp (:116-178), a
// different hash: the XOR of one FNV pair hash per entry, mixed into the
// length seed.
void duplicate_row_hashes(const CsrV& A, const uint8_t* keep_row, double ctol, uint64_t* h) {
  const int32_t m = A.rows;
  const bool wide = static_cast<int64_t>(A.nnz) > 32LL * static_cast<int64_t>(m);
  for (int32_t i = 0; i < m; ++i) {
    if (!keep_row[i]) { h[i] = static_cast<uint64_t>(i); continue; }
    const int32_t b = A.ptr[i], e = A.ptr[i + 1];
    if (e - b <= 0 || std::fabs(A.val[b]) <= ctol) { h[i] = fnv_mix(FNV_OFFSET, static_cast<uint64_t>(i)); continue; }
    const double piv = A.val[b];
    if (wide) {
      uint64_t agg = 0;
      for (int32_t q = b; q < e; ++q)
        agg ^= fnv_mix(fnv_mix(FNV_OFFSET, static_cast<uint64_t>(static_cast<int64_t>(A.idx[q]))),
                       dbits(A.val[q] / piv));
      h[i] = fnv_mix(fnv_mix(FNV_OFFSET, static_cast<uint64_t>(e - b)), agg);
      continue;
    }
    uint64_t x = fnv_mix(FNV_OFFSET, static_cast<uint64_t>(e - b));
    for (int32_t q = b; q < e; ++q) {
      x = fnv_mix(x, static_cast<uint64_t>(static_cast<int64_t>(A.idx[q])));
      x = fnv_mix(x, dbits(A.val[q] / piv));
    }
    h[i] = x;
  }
}

// ---------------------------------------------------------------------------
// rules/rule_duplicate_rows.cu:78-613
void rule_duplicate_rows(Plan& p, const LpV& lp, const Stats&, const Params& prm) {
  if (p.has_infeasible || p.has_unbounded) return;
  const int32_t m = lp.A.rows;
  if (m <= 1) return;
  const double ctol = prm.zero_tol, itol = prm.feasibility_tol;
  const CsrV& A = lp.A;
  std::vector<uint64_t> h(m);
  duplicate_row_hashes(A, p.keep_row, ctol, h.data());
  std::vector<int32_t> order(m);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    return h[a] != h[b] ? h[a] < h[b] : a < b;
  });
  // _duplicate_row_ratio (:180-219)
  auto ratio_of = [&](int32_t i, int32_t k, double* r) -> bool {
    const int32_t bi = A.ptr[i], li = A.ptr[i + 1] - bi;
    const int32_t bk = A.ptr[k], lk = A.ptr[k + 1] - bk;
    if (li != lk) return false;
    if (li <= 0) { *r = 1.0; return true; }
    double ratio = 0.0;
    bool set = false;
    for (int32_t o = 0; o < li; ++o) {
      if (A.idx[bi + o] != A.idx[bk + o]) return false;
      const double ai = A.val[bi + o], ak = A.val[bk + o];
      if (!set) {
        if (std::fabs(ak) <= ctol) return false;
        ratio = ai / ak;
        set = true;
      }
      if (std::fabs(ai - ratio * ak) > ctol) return false;
    }
    *r = ratio;
    return set;
  };
  auto scaled = [](double lo, double hi, double r, double* sl, double* su) {
    const double a = lo * r, b = hi * r;
    if (r >= 0.0) { *sl = a; *su = b; } else { *sl = b; *su = a; }
  };
  struct Rec { int32_t step, kept, deleted; double ratio, kal, kau, dal, dau; };
  std::vector<Rec> recs;
  std::vector<uint8_t> del(m, 0);
  bool infeasible = false;
  // _kernel_duplicate_row_groups (:243-381), one bucket at a time
  for (int32_t s0 = 0; s0 < m;) {
    int32_t e0 = s0;
    while (e0 + 1 < m && h[order[e0 + 1]] == h[order[s0]]) ++e0;
    const int32_t s = s0, e = e0;
    s0 = e0 + 1;
    if (e <= s) continue;
    int32_t seed = -1;
    bool seed_eq = false;
    for (int32_t a = s; a <= e; ++a) {
      const int32_t row = order[a];
      if (!p.keep_row[row]) continue;
      const bool eq = fin(p.AL[row]) && fin(p.AU[row]) && std::fabs(p.AL[row] - p.AU[row]) <= itol;
      if (seed < 0 || (eq && !seed_eq)) { seed = row; seed_eq = eq; }
    }
    if (seed < 0) continue;
    int32_t rep = seed;
    double rsl = p.AL[seed], rsu = p.AU[seed];
    for (int32_t a = s; a <= e; ++a) {
      const int32_t row = order[a];
      if (row == seed || !p.keep_row[row]) continue;
      double r;
      if (!ratio_of(seed, row, &r)) continue;
      double sl, su;
      scaled(p.AL[row], p.AU[row], r, &sl, &su);
      if (rsl <= sl + itol && su <= rsu + itol) { rep = row; rsl = sl; rsu = su; }
    }
    double rl = p.AL[rep], ru = p.AU[rep];
    bool peer = false, bad = false;
    int32_t step = 0;
    std::vector<Rec> group_recs;
    std::vector<int32_t> group_del;
    for (int32_t a = s; a <= e; ++a) {
      const int32_t row = order[a];
      if (row == rep || !p.keep_row[row]) continue;
      double r;
      if (!ratio_of(rep, row, &r)) continue;
      peer = true;
      double sl, su;
      scaled(p.AL[row], p.AU[row], r, &sl, &su);
      if (std::fmax(rl, sl) > std::fmin(ru, su) + itol) { bad = true; break; }
      ++step;
      group_recs.push_back(Rec{step, rep, row, r, rl, ru, p.AL[row], p.AU[row]});
      rl = std::fmax(rl, sl);
      ru = std::fmin(ru, su);
      group_del.push_back(row);
    }
    if (bad) { infeasible = true; continue; }
    for (int32_t row : group_del) del[row] = 1;
    recs.insert(recs.end(), group_recs.begin(), group_recs.end());
    if (peer) { p.AL[rep] = rl; p.AU[rep] = ru; }
  }
  if (infeasible) { p.has_infeasible = true; return; }
  bool changed = false;
  for (int32_t i = 0; i < m; ++i) if (del[i]) { p.keep_row[i] = 0; changed = true; }
  if (!changed) return;
  if (prm.record_postsolve_tape) {
    std::sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) {
      return a.step != b.step ? a.step < b.step : a.deleted < b.deleted;
    });
    for (const Rec& r : recs)
      p.tape.append(DuplicateRow, {r.kept, r.deleted}, {r.ratio, r.kal, r.kau, r.dal, r.dau}, DualExact);
  }
  p.has_row_action = true;
  p.has_change = true;
}

// ---------------------------------------------------------------------------
// rules/rule_infeasible_fixed_variables.cu:42-266
void rule_infeasible_fixed_variables(Plan& p, const LpV& lp, const Params& prm) {
  if (p.has_infeasible || p.has_unbounded) return;
  const int32_t n = lp.A.cols, m = lp.A.rows;
  if (n == 0) return;
  std::vector<uint8_t> fixed(n, 0);
  std::vector<double> v(n, 0.0), contrib(n, 0.0);
  int32_t count = 0;
  for (int32_t j = 0; j < n; ++j) {
    if (!p.keep_col[j]) continue;
    const double lj = p.l[j], uj = p.u[j];
    if (fin(lj) && fin(uj) && std::fabs(uj - lj) <= prm.bound_tol) {
      v[j] = lj + 0.5 * (uj - lj);
      fixed[j] = 1;
      contrib[j] = p.c[j] * v[j];
      ++count;
    }
  }
  if (count == 0) return;
  if (prm.record_postsolve_tape && prm.record_postsolve_tape_cpu) {
    for (int32_t j = 0; j < n; ++j) {
      if (!fixed[j]) continue;
      std::vector<int32_t> idx{j};
      std::vector<double> vals{v[j], p.c[j]};
      for (int32_t q = lp.AT.ptr[j]; q < lp.AT.ptr[j + 1]; ++q) {
        if (!p.keep_row[lp.AT.idx[q]]) continue;
        idx.push_back(lp.AT.idx[q]);
        vals.push_back(lp.AT.val[q]);
      }
      p.tape.append(FixedCol, idx, vals, DualMinimal);
    }
  }
  // Row shift over every row (the reference does not test keep_row here).
  std::vector<double> shift(m, 0.0);
  for (int32_t i = 0; i < m; ++i)
    for (int32_t q = lp.A.ptr[i]; q < lp.A.ptr[i + 1]; ++q)
      if (fixed[lp.A.idx[q]]) shift[i] += lp.A.val[q] * v[lp.A.idx[q]];
  for (int32_t j = 0; j < n; ++j) {
    if (!fixed[j]) continue;
    p.l[j] = v[j];
    p.u[j] = v[j];
    p.keep_col[j] = 0;
  }
  for (int32_t i = 0; i < m; ++i) { p.AL[i] -= shift[i]; p.AU[i] -= shift[i]; }
  p.obj_delta += ordered_sum(contrib);
  p.has_col_action = true;
  p.has_change = true;
}

// rules/rule_empty_cols.cu:42-198
void rule_empty_cols(Plan& p, const LpV& lp, const Stats& s, const Params& prm) {
  if (p.has_infeasible || p.has_unbounded) return;
  const int32_t n = lp.A.cols;
  if (n == 0) return;
  std::vector<uint8_t> fixed(n, 0);
  std::vector<double> v(n, 0.0), contrib(n, 0.0);
  bool infeasible = false, unbounded = false, any = false;
  for (int32_t j = 0; j < n; ++j) {
    if (!(p.keep_col[j] && s.empty_col_mask[j])) continue;
    const double cj = p.c[j], lj = p.l[j], uj = p.u[j];
    if (lj > uj + prm.bound_tol) { infeasible = true; continue; }
    if ((cj > prm.zero_tol && !fin(lj)) || (cj < -prm.zero_tol && !fin(uj))) { unbounded = true; continue; }
    double val = 0.0;
    if (cj > prm.zero_tol) val = lj;
    else if (cj < -prm.zero_tol) val = uj;
    else if (fin(lj)) val = lj;
    else if (fin(uj)) val = uj;
    fixed[j] = 1;
    v[j] = val;
    contrib[j] = cj * val;
    any = true;
  }
  if (infeasible) { p.has_infeasible = true; return; }
  if (unbounded) { p.has_unbounded = true; return; }
  if (!any) return;
  if (prm.record_postsolve_tape && prm.record_postsolve_tape_cpu)
    for (int32_t j = 0; j < n; ++j)
      if (fixed[j]) p.tape.append(FixedCol, {j}, {v[j], p.c[j]}, DualMinimal);
  for (int32_t j = 0; j < n; ++j) {
    if (!fixed[j]) continue;
    p.l[j] = v[j];
    p.u[j] = v[j];
    p.keep_col[j] = 0;
  }
  p.obj_delta += ordered_sum(contrib);
  p.has_col_action = true;
  p.has_change = true;
}

// ---------------------------------------------------------------------------
// rules/rule_column_singletons.cu
namespace {
constexpr uint8_t INEQ_NONE = 0, INEQ_ELIM = 1, INEQ_LHS_TO_RHS = 2, INEQ_RHS_TO_LHS = 3;

// :448-458
bool cs_direct_unbounded(double cj, double a, double lhs, double rhs, double lb, double ub, double zt) {
  return (cj > zt && a > zt && !fin(lhs) && !fin(lb)) || (cj > zt && a < -zt && !fin(rhs) && !fin(lb)) ||
         (cj < -zt && a < -zt && !fin(lhs) && !fin(ub)) || (cj < -zt && a > zt && !fin(rhs) && !fin(ub));
}
// :717-751
bool cs_free_above(double a, double lhs, double rhs, double ub, double rmin, double rmax, double tol) {
  if (!fin(ub)) return true;
  double iu = INF;
  if (a > 0.0 && fin(rhs)) iu = (rhs - rmin) / a;
  else if (a < 0.0 && fin(lhs)) iu = (lhs - rmax) / a;
  return iu <= ub + tol;
}
bool cs_free_below(double a, double lhs, double rhs, double lb, double rmin, double rmax, double tol) {
  if (!fin(lb)) return true;
  double il = -INF;
  if (a > 0.0 && fin(lhs)) il = (lhs - rmax) / a;
  else if (a < 0.0 && fin(rhs)) il = (rhs - rmin) / a;
  return il >= lb - tol;
}
// :753-800
double cs_active_side(double cj, double a, double lhs, double rhs, double zt) {
  if ((cj > zt && a > 0.0) || (cj < -zt && a < 0.0)) return lhs;
  if ((cj > zt && a < 0.0) || (cj < -zt && a > 0.0)) return rhs;
  return fin(lhs) ? lhs : rhs;
}
uint8_t cs_ineq_action(double cj, double a, double lhs, double rhs, bool above, bool below, double zt,
                       double* side) {
  if (above && below) {
    *side = cs_active_side(cj, a, lhs, rhs, zt);
    return fin(*side) ? INEQ_ELIM : INEQ_NONE;
  }
  if (((cj < -zt && a > 0.0 && above) || (cj > zt && a < 0.0 && below)) && fin(rhs)) {
    *side = rhs;
    return INEQ_LHS_TO_RHS;
  }
  if (((cj > zt && a > 0.0 && below) || (cj < -zt && a < 0.0 && above)) && fin(lhs)) {
    *side = lhs;
    return INEQ_RHS_TO_LHS;
  }
  *side = 0.0;
  return INEQ_NONE;
}

struct RowCache {
  std::vector<double> fmin, fmax;
  std::vector<int32_t> mn, mp, xn, xp;
  std::vector<uint8_t> invalid;
};

// :653-713
bool cs_rest_bounds(const RowCache& rc, int32_t row, int32_t col, double a, const double* l, const double* u,
                    double* rest_min, double* rest_max) {
  if (rc.invalid[row]) return false;
  const double tmin = a >= 0.0 ? a * l[col] : a * u[col];
  const double tmax = a >= 0.0 ? a * u[col] : a * l[col];
  if (std::isnan(tmin) || std::isnan(tmax)) return false;
  int32_t mn = rc.mn[row], mp = rc.mp[row], xn = rc.xn[row], xp = rc.xp[row];
  double vmin = rc.fmin[row], vmax = rc.fmax[row];
  if (std::isinf(tmin)) { if (tmin < 0.0) --mn; else --mp; } else vmin -= tmin;
  if (std::isinf(tmax)) { if (tmax < 0.0) --xn; else --xp; } else vmax -= tmax;
  if (mn < 0 || mp < 0 || xn < 0 || xp < 0 || (mn > 0 && mp > 0) || (xn > 0 && xp > 0)) return false;
  *rest_min = mn > 0 ? -INF : (mp > 0 ? INF : vmin);
  *rest_max = xn > 0 ? -INF : (xp > 0 ? INF : vmax);
  return !std::isnan(*rest_min) && !std::isnan(*rest_max);
}
}  // namespace

// rules/rule_column_singletons.cu:1631-1928
void rule_column_singletons(Plan& p, const LpV& lp, const Stats& s, const Params& prm, bool eq_mode,
                            bool skip_precheck) {
  if (p.has_infeasible || p.has_unbounded) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0 || n == 0) return;
  const double tol = prm.bound_tol, zt = prm.zero_tol;
  if (!skip_precheck) {
    int32_t cnt = 0, eq = 0;
    for (int32_t j = 0; j < n; ++j) {
      if (!p.keep_col[j] || !s.column_singleton_mask[j]) continue;
      const int32_t row = s.column_singleton_row[j];
      if (row >= 0 && p.keep_row[row]) {
        ++cnt;
        if (fin(p.AL[row]) && fin(p.AU[row]) && std::fabs(p.AU[row] - p.AL[row]) <= tol) ++eq;
      }
    }
    if (cnt == 0 || (eq_mode && eq == 0) || (!eq_mode && cnt == eq)) return;
  }
  // Activity cache over rows that hold a singleton column (:455-535)
  std::vector<uint8_t> needed(m, 0);
  for (int32_t j = 0; j < n; ++j) {
    if (!p.keep_col[j] || !s.column_singleton_mask[j]) continue;
    const int32_t row = s.column_singleton_row[j];
    if (row >= 0 && row < m && p.keep_row[row]) needed[row] = 1;
  }
  RowCache rc;
  rc.fmin.assign(m, 0.0); rc.fmax.assign(m, 0.0);
  rc.mn.assign(m, 0); rc.mp.assign(m, 0); rc.xn.assign(m, 0); rc.xp.assign(m, 0);
  rc.invalid.assign(m, 0);
  for (int32_t i = 0; i < m; ++i) {
    if (!needed[i] || !p.keep_row[i]) continue;
    for (int32_t q = lp.A.ptr[i]; q < lp.A.ptr[i + 1]; ++q) {
      const int32_t col = lp.A.idx[q];
      if (!p.keep_col[col]) continue;
      const double a = lp.A.val[q];
      const double tmin = a >= 0.0 ? a * p.l[col] : a * p.u[col];
      const double tmax = a >= 0.0 ? a * p.u[col] : a * p.l[col];
      if (std::isnan(tmin) || std::isnan(tmax)) { rc.invalid[i] = 1; continue; }
      if (std::isinf(tmin)) { if (tmin < 0.0) ++rc.mn[i]; else ++rc.mp[i]; } else rc.fmin[i] += tmin;
      if (std::isinf(tmax)) { if (tmax < 0.0) ++rc.xn[i]; else ++rc.xp[i]; } else rc.fmax[i] += tmax;
    }
  }
  // Row owner = smallest qualifying singleton column (:802-921)
  std::vector<int32_t> owner(m, INT_MAX);
  bool unbounded = false;
  for (int32_t j = 0; j < n; ++j) {
    if (!p.keep_col[j] || !s.column_singleton_mask[j]) continue;
    const int32_t row = s.column_singleton_row[j];
    if (row < 0 || row >= m || !p.keep_row[row]) continue;
    const double a = s.column_singleton_val[j];
    if (std::fabs(a) <= zt) continue;
    int32_t live = 0;
    for (int32_t q = lp.A.ptr[row]; q < lp.A.ptr[row + 1]; ++q) {
      if (p.keep_col[lp.A.idx[q]]) { if (++live > 1) break; }
    }
    if (live <= 1) continue;
    const double lhs = p.AL[row], rhs = p.AU[row], cj = p.c[j];
    if (!eq_mode && cs_direct_unbounded(cj, a, lhs, rhs, p.l[j], p.u[j], zt)) { unbounded = true; continue; }
    double rmin = 0.0, rmax = 0.0;
    if (!cs_rest_bounds(rc, row, j, a, p.l, p.u, &rmin, &rmax)) continue;
    const bool is_eq = fin(lhs) && fin(rhs) && std::fabs(lhs - rhs) <= tol;
    if (is_eq) {
      if (!eq_mode) continue;
      const double x1 = (rhs - rmin) / a, x2 = (rhs - rmax) / a;
      const double ilb = std::fmin(x1, x2), iub = std::fmax(x1, x2);
      const bool above = !fin(p.u[j]) || iub <= p.u[j] + tol;
      const bool below = !fin(p.l[j]) || ilb >= p.l[j] - tol;
      if (!(above || below)) continue;
    } else {
      if (eq_mode) continue;
      const bool above = cs_free_above(a, lhs, rhs, p.u[j], rmin, rmax, tol);
      const bool below = cs_free_below(a, lhs, rhs, p.l[j], rmin, rmax, tol);
      double side;
      if (cs_ineq_action(cj, a, lhs, rhs, above, below, zt, &side) == INEQ_NONE) continue;
    }
    owner[row] = std::min(owner[row], j);
  }
  if (unbounded) { p.has_unbounded = true; return; }
  // _kernel_process_column_singletons (:923-1061)
  std::vector<uint8_t> row_del(m, 0), lhs_ch(m, 0), rhs_ch(m, 0), col_del(n, 0);
  std::vector<double> side_of(n, 0.0), contrib(n, 0.0);
  std::vector<double> AL_new(p.AL, p.AL + m), AU_new(p.AU, p.AU + m);
  bool st_eq = false, st_dual = false;
  for (int32_t j = 0; j < n; ++j) {
    if (!p.keep_col[j] || !s.column_singleton_mask[j]) continue;
    const int32_t row = s.column_singleton_row[j];
    if (row < 0 || row >= m || !p.keep_row[row] || owner[row] != j) continue;
    const double a = s.column_singleton_val[j];
    if (std::fabs(a) <= zt) continue;
    const double lhs = AL_new[row], rhs = AU_new[row], cj = p.c[j];
    if (!eq_mode && cs_direct_unbounded(cj, a, lhs, rhs, p.l[j], p.u[j], zt)) { unbounded = true; continue; }
    double rmin = 0.0, rmax = 0.0;
    if (!cs_rest_bounds(rc, row, j, a, p.l, p.u, &rmin, &rmax)) continue;
    const bool is_eq = fin(lhs) && fin(rhs) && std::fabs(lhs - rhs) <= tol;
    if (is_eq) {
      if (!eq_mode) continue;
      const double x1 = (rhs - rmin) / a, x2 = (rhs - rmax) / a;
      const double ilb = std::fmin(x1, x2), iub = std::fmax(x1, x2);
      const bool above = !fin(p.u[j]) || iub <= p.u[j] + tol;
      const bool below = !fin(p.l[j]) || ilb >= p.l[j] - tol;
      if (!(above || below)) continue;
      side_of[j] = rhs;
      contrib[j] = cj * rhs / a;
      if (above && below) {
        row_del[row] = 1;
      } else {
        // _apply_column_singleton_eq_one_sided_row_update (:480-494)
        const double bval = above ? p.l[j] : p.u[j];
        const bool keep_lower = above ? (a < 0.0) : (a > 0.0);
        const double shifted = AU_new[row] - a * bval;
        if (keep_lower) { AL_new[row] = shifted; AU_new[row] = INF; }
        else { AL_new[row] = -INF; AU_new[row] = shifted; }
      }
      col_del[j] = 1;
      st_eq = true;
      continue;
    }
    if (eq_mode) continue;
    const bool above = cs_free_above(a, lhs, rhs, p.u[j], rmin, rmax, tol);
    const bool below = cs_free_below(a, lhs, rhs, p.l[j], rmin, rmax, tol);
    double side = 0.0;
    const uint8_t act = cs_ineq_action(cj, a, lhs, rhs, above, below, zt, &side);
    if (act == INEQ_NONE) continue;
    if (act == INEQ_ELIM) {
      if (!fin(lhs) || std::fabs(lhs - side) > tol) lhs_ch[row] = 1;
      if (!fin(rhs) || std::fabs(rhs - side) > tol) rhs_ch[row] = 1;
      AL_new[row] = side;
      AU_new[row] = side;
    } else if (act == INEQ_LHS_TO_RHS) {
      lhs_ch[row] = 1;
      AL_new[row] = rhs;
    } else {
      rhs_ch[row] = 1;
      AU_new[row] = lhs;
    }
    st_dual = true;
  }
  if (unbounded) { p.has_unbounded = true; return; }
  if (!(st_eq || st_dual)) return;
  std::vector<double> c_new(p.c, p.c + n);
  if (st_eq && eq_mode) {
    // _kernel_apply_singleton_objective_updates_deterministic (:1063-1091)
    for (int32_t col = 0; col < n; ++col) {
      if (!p.keep_col[col] || col_del[col]) continue;
      double upd = c_new[col];
      for (int32_t q = lp.AT.ptr[col]; q < lp.AT.ptr[col + 1]; ++q) {
        const int32_t ow = owner[lp.AT.idx[q]];
        if (ow < 0 || ow >= n || !col_del[ow]) continue;
        upd += -(p.c[ow] * lp.AT.val[q] / s.column_singleton_val[ow]);
      }
      c_new[col] = upd;
    }
  }
  if (st_eq && eq_mode && prm.record_postsolve_tape) {
    // append_subcol_tape_compacted_from_device (:1358-1629): one SubCol per
    // owning row in ascending row order.
    for (int32_t row = 0; row < m; ++row) {
      const int32_t col = owner[row];
      if (col < 0 || col >= n || !col_del[col]) continue;
      double pivot = 0.0;
      std::vector<int32_t> sc;
      std::vector<double> sv;
      for (int32_t q = lp.A.ptr[row]; q < lp.A.ptr[row + 1]; ++q) {
        const int32_t rc2 = lp.A.idx[q];
        if (rc2 == col) pivot = lp.A.val[q];
        else if (p.keep_col[rc2]) { sc.push_back(rc2); sv.push_back(lp.A.val[q]); }
      }
      if (std::fabs(pivot) <= zt) continue;
      std::vector<int32_t> idx{col, row, static_cast<int32_t>(sc.size())};
      idx.insert(idx.end(), sc.begin(), sc.end());
      std::vector<double> v{pivot, side_of[col], p.l[col], p.u[col], p.c[col], row_del[row] ? 1.0 : 0.0};
      v.insert(v.end(), sv.begin(), sv.end());
      p.tape.append(SubCol, idx, v, DualMinimal);
    }
  }
  if (st_dual && !eq_mode && prm.record_postsolve_tape) {
    for (int32_t row = 0; row < m; ++row)
      if (lhs_ch[row] || rhs_ch[row]) p.tape.append(EqToIneq, {row}, {0.0}, DualMinimal);
  }
  for (int32_t i = 0; i < m; ++i) {
    if (row_del[i]) p.keep_row[i] = 0;
    p.AL[i] = AL_new[i];
    p.AU[i] = AU_new[i];
  }
  for (int32_t j = 0; j < n; ++j) {
    if (col_del[j]) p.keep_col[j] = 0;
    p.c[j] = c_new[j];
  }
  p.obj_delta += ordered_sum(contrib);
  p.has_row_action = true;
  p.has_col_action = p.has_col_action || st_eq;
  p.has_change = true;
}

// ---------------------------------------------------------------------------
// rules/rule_doubleton_equations.cu
namespace {
constexpr int DBL_MATCH_ROUNDS = 8;
constexpr int DBL_INNER_ROUNDS = 1024;
constexpr double DBL_MAX_RATIO = 1.0e3;

// :143-153
bool dbl_integral_ratio(double num, double den, double tol) {
  if (std::fabs(den) <= tol) return false;
  const double r = std::fabs(num / den);
  if (!fin(r)) return false;
  return std::fabs(r - std::nearbyint(r)) <= tol;
}
// :155-162
bool dbl_pivot_ok(double keep_val, double elim_val) {
  if (std::fabs(elim_val) <= 0.0) return false;
  const double r = std::fabs(keep_val / elim_val);
  return r <= DBL_MAX_RATIO && r >= 1.0 / DBL_MAX_RATIO;
}
// :164-207
void dbl_choose(int32_t c1, double v1, int32_t c2, double v2, const std::vector<int32_t>& nnz, double tol,
                int32_t* ec, double* ev, int32_t* kc, double* kv) {
  const bool i12 = dbl_integral_ratio(v1, v2, tol), i21 = dbl_integral_ratio(v2, v1, tol);
  bool elim_first;
  if (nnz[c1] == 1 && nnz[c2] != 1) elim_first = true;
  else if (nnz[c1] != 1 && nnz[c2] == 1) elim_first = false;
  else if (i12 && !i21) elim_first = false;
  else if (i21 && !i12) elim_first = true;
  else if (nnz[c1] < nnz[c2]) elim_first = true;
  else elim_first = false;
  if (elim_first) { *ec = c1; *ev = v1; *kc = c2; *kv = v2; }
  else { *ec = c2; *ev = v2; *kc = c1; *kv = v1; }
}
// :209-236 (structural fill-in proxy over sorted AT rows)
int32_t dbl_fill_in(const CsrV& AT, int32_t keep, int32_t elim) {
  int32_t fill = -1;
  int32_t jj = AT.ptr[keep], ks = AT.ptr[keep + 1], kk = AT.ptr[elim], es = AT.ptr[elim + 1];
  if (ks <= jj || es <= kk) return fill;
  while (jj < ks && kk < es) {
    const int32_t kr = AT.idx[jj], er = AT.idx[kk];
    if (kr == er) { ++jj; ++kk; }
    else if (er < kr) { ++kk; ++fill; }
    else ++jj;
  }
  fill += es - kk;
  return fill;
}
// :676-744 (live variant)
int32_t dbl_next_live(const CsrV& AT, const uint8_t* keep_row, double zt, int32_t col, int32_t* pos) {
  const int32_t stop = AT.ptr[col + 1];
  while (*pos < stop) {
    const int32_t r = AT.idx[*pos];
    if (keep_row[r] && std::fabs(AT.val[*pos]) > zt) return r;
    ++(*pos);
  }
  return INT_MAX;
}
int32_t dbl_fill_in_live(const CsrV& AT, const uint8_t* keep_row, double zt, int32_t keep, int32_t elim) {
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
// Live two-entry equality row -> (elim, keep) candidate (:746-805, :389-437)
bool dbl_row_candidate(const CsrV& A, const Plan& p, int32_t i, const std::vector<int32_t>& nnz,
                       double zt, double tol, double rtol, int32_t* ec, double* ev, int32_t* kc, double* kv) {
  if (!p.keep_row[i]) return false;
  const double lhs = p.AL[i], rhs = p.AU[i];
  if (!fin(lhs) || !fin(rhs) || std::fabs(lhs - rhs) > tol) return false;
  int32_t live = 0, c1 = -1, c2 = -1;
  double v1 = 0.0, v2 = 0.0;
  for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
    const int32_t col = A.idx[q];
    const double a = A.val[q];
    if (p.keep_col[col] && std::fabs(a) > zt) {
      ++live;
      if (live == 1) { c1 = col; v1 = a; }
      else if (live == 2) { c2 = col; v2 = a; }
      else return false;
    }
  }
  if (live != 2) return false;
  dbl_choose(c1, v1, c2, v2, nnz, rtol, ec, ev, kc, kv);
  return dbl_pivot_ok(*kv, *ev);
}

struct DblRec {
  int32_t elim, keep, row;
  double ev, kv, au, el, eu, kl, ku, nkl, nku, ec;
};

// One batch (:1432-1677).  Returns true when the plan changed.
bool dbl_batch_once(Plan& p, const LpV& lp, const Csr& AT_src_own, const CsrV& AT_src, const Params& prm,
                    std::vector<DblRec>* tape_recs, double* obj_subst, double* obj_fixed) {
  (void)AT_src_own;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0 || n == 0) return false;
  const CsrV A = p.has_new_A ? p.new_A.view() : lp.A;
  const double zt = prm.zero_tol, tol = prm.bound_tol;
  const double rtol = std::fmax(1.0e-9, prm.bound_tol);
  std::vector<int32_t> nnz(n, 0);
  for (int32_t col = 0; col < n; ++col) {
    if (!p.keep_col[col]) continue;
    int32_t cnt = 0;
    for (int32_t q = AT_src.ptr[col]; q < AT_src.ptr[col + 1]; ++q)
      if (p.keep_row[AT_src.idx[q]] && std::fabs(AT_src.val[q]) > zt) ++cnt;
    nnz[col] = cnt;
  }
  std::vector<uint8_t> acceptable(m, 0);
  std::vector<int32_t> cec(m, -1), ckc(m, -1);
  std::vector<double> cev(m, 0.0), ckv(m, 0.0);
  int32_t acceptable_count = 0;
  for (int32_t i = 0; i < m; ++i) {
    int32_t ec, kc;
    double ev, kv;
    if (!dbl_row_candidate(A, p, i, nnz, zt, tol, rtol, &ec, &ev, &kc, &kv)) continue;
    cec[i] = ec; cev[i] = ev; ckc[i] = kc; ckv[i] = kv;
    if (nnz[ec] > nnz[kc] + prm.doubleton_equations_max_fill_in_proxy + 1) continue;
    if (dbl_fill_in_live(AT_src, p.keep_row, zt, kc, ec) <= prm.doubleton_equations_max_fill_in_proxy) {
      acceptable[i] = 1;
      ++acceptable_count;
    }
  }
  // Greedy matching: 8 rounds of min-row-index claims (:876-913)
  std::vector<uint8_t> active(acceptable), selected(m, 0), blocked(n, 0);
  std::vector<int32_t> owner(n);
  for (int r = 0; r < DBL_MATCH_ROUNDS; ++r) {
    std::fill(owner.begin(), owner.end(), INT_MAX);
    for (int32_t i = 0; i < m; ++i) {
      if (!active[i] || blocked[ckc[i]] || blocked[cec[i]]) continue;
      owner[ckc[i]] = std::min(owner[ckc[i]], i);
      owner[cec[i]] = std::min(owner[cec[i]], i);
    }
    for (int32_t i = 0; i < m; ++i) {
      if (!active[i]) continue;
      const int32_t k = ckc[i], e = cec[i];
      if (!blocked[k] && !blocked[e] && owner[k] == i && owner[e] == i) {
        selected[i] = 1; active[i] = 0; blocked[k] = 1; blocked[e] = 1;
      }
    }
  }
  std::vector<int32_t> sel;
  for (int32_t i = 0; i < m; ++i) if (selected[i]) sel.push_back(i);
  const int32_t k = static_cast<int32_t>(sel.size());
  const double ratio = acceptable_count == 0 ? 0.0 : static_cast<double>(k) / acceptable_count;
  if (k == 0 || (prm.doubleton_equations_scan &&
                 (k < prm.doubleton_equations_min_selected_per_batch ||
                  ratio < prm.doubleton_equations_min_selected_ratio)))
    return false;
  std::vector<DblRec> recs(k);
  for (int32_t t = 0; t < k; ++t) {
    const int32_t row = sel[t], kc = ckc[row], ec = cec[row];
    recs[t] = DblRec{ec, kc, row, cev[row], ckv[row], p.AU[row], p.l[ec], p.u[ec], p.l[kc], p.u[kc],
                     0.0, 0.0, p.c[ec]};
  }
  // _kernel_apply_doubleton_equations_batch (:916-1004)
  std::vector<uint8_t> subst(k, 0), fixedp(k, 0);
  std::vector<double> alpha(k), beta(k), fixed_at(k, 0.0);
  bool infeasible = false;
  std::vector<double> osub, ofix;
  for (int32_t t = 0; t < k; ++t) {
    const int32_t row = sel[t], kc = ckc[row], ec = cec[row];
    const double kv = ckv[row], ev = cev[row], rhs = p.AU[row];
    const double al = -kv / ev, be = rhs / ev;
    alpha[t] = al; beta[t] = be;
    double ml = -INF, mu = INF;
    if (fin(p.l[ec])) { const double b = (p.l[ec] - be) / al; if (al > 0.0) ml = b; else mu = b; }
    if (fin(p.u[ec])) { const double b = (p.u[ec] - be) / al; if (al > 0.0) mu = b; else ml = b; }
    if (ml > mu) std::swap(ml, mu);
    const double nkl = fin(ml) ? std::fmax(p.l[kc], ml) : p.l[kc];
    const double nku = fin(mu) ? std::fmin(p.u[kc], mu) : p.u[kc];
    if (nkl > nku + tol) { infeasible = true; continue; }
    const bool keep_fixed = fin(nkl) && fin(nku) && nku <= nkl + tol;
    if (keep_fixed) {
      const double fat = 0.5 * (nkl + nku);
      p.l[kc] = fat; p.u[kc] = fat;
      p.keep_col[kc] = 0;
      fixedp[t] = 1;
      fixed_at[t] = fat;
      ofix.push_back(p.c[kc] * fat);
    } else {
      p.l[kc] = nkl; p.u[kc] = nku;
      p.c[kc] += p.c[ec] * al;
      p.keep_row[row] = 0;
      p.keep_col[ec] = 0;
      subst[t] = 1;
      osub.push_back(p.c[ec] * be);
    }
  }
  *obj_subst = ordered_sum(osub);
  *obj_fixed = ordered_sum(ofix);
  if (infeasible) { p.has_infeasible = true; return false; }
  int32_t subst_count = 0, fixed_count = 0;
  for (int32_t t = 0; t < k; ++t) { subst_count += subst[t]; fixed_count += fixedp[t]; }
  for (int32_t t = 0; t < k; ++t) {
    if (!subst[t]) continue;
    DblRec r = recs[t];
    r.nkl = p.l[r.keep];
    r.nku = p.u[r.keep];
    if (prm.record_postsolve_tape) tape_recs->push_back(r);
  }
  // Per-column lookups for the row shift and the delta entries.
  std::vector<double> fixed_val_of(n, 0.0), beta_of(n, 0.0), alpha_of(n, 0.0);
  std::vector<uint8_t> is_fixed_keep(n, 0), is_elim(n, 0);
  std::vector<int32_t> row_of(n, -1), keep_of(n, -1);
  for (int32_t t = 0; t < k; ++t) {
    const int32_t row = sel[t];
    if (fixedp[t]) { is_fixed_keep[ckc[row]] = 1; fixed_val_of[ckc[row]] = p.l[ckc[row]]; }
    if (subst[t]) {
      const int32_t ec = cec[row];
      is_elim[ec] = 1; beta_of[ec] = beta[t]; alpha_of[ec] = alpha[t];
      row_of[ec] = row; keep_of[ec] = ckc[row];
    }
  }
  // Row shift: fixed-pair keep columns first, then substituted columns, each
  // in ascending column order along the row (:1040-1076).
  std::vector<double> shift(m, 0.0);
  if (fixed_count > 0) {
    for (int32_t i = 0; i < m; ++i) {
      if (!p.keep_row[i]) continue;
      for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
        const int32_t j = A.idx[q];
        if (is_fixed_keep[j] && std::fabs(A.val[q]) > zt) shift[i] += A.val[q] * fixed_val_of[j];
      }
    }
  }
  if (subst_count > 0) {
    // build_doubleton_batch_new_A (:1184-1314)
    int64_t total = 0;
    std::vector<std::vector<std::pair<int32_t, double>>> delta(m);
    for (int32_t i = 0; i < m; ++i) {
      if (!p.keep_row[i]) continue;
      for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
        const int32_t j = A.idx[q];
        if (!is_elim[j] || row_of[j] == i || !(std::fabs(A.val[q]) > zt)) continue;
        delta[i].emplace_back(keep_of[j], alpha_of[j] * A.val[q]);
        ++total;
      }
    }
    if (total > 0) {
      for (int32_t i = 0; i < m; ++i) {
        if (!p.keep_row[i]) continue;
        for (int32_t q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
          const int32_t j = A.idx[q];
          if (is_elim[j] && row_of[j] != i && std::fabs(A.val[q]) > zt) shift[i] += A.val[q] * beta_of[j];
        }
      }
      Csr out;
      out.rows = A.rows;
      out.cols = A.cols;
      out.ptr.assign(1, 0);
      for (int32_t i = 0; i < m; ++i) {
        auto& d = delta[i];
        std::sort(d.begin(), d.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        int32_t q = A.ptr[i];
        const int32_t qe = A.ptr[i + 1];
        size_t r = 0;
        while (q < qe || r < d.size()) {
          const int32_t sc = q < qe ? A.idx[q] : INT_MAX;
          const int32_t dc = r < d.size() ? d[r].first : INT_MAX;
          const int32_t nc = std::min(sc, dc);
          double merged = 0.0;
          while (q < qe && A.idx[q] == nc) { merged += A.val[q]; ++q; }
          while (r < d.size() && d[r].first == nc) { merged += d[r].second; ++r; }
          if (std::fabs(merged) > zt) { out.idx.push_back(nc); out.val.push_back(merged); }
        }
        out.ptr.push_back(static_cast<int32_t>(out.idx.size()));
      }
      p.new_A = std::move(out);
      p.has_new_A = true;
    }
  }
  for (int32_t i = 0; i < m; ++i) { p.AL[i] -= shift[i]; p.AU[i] -= shift[i]; }
  p.has_row_action = subst_count > 0;
  p.has_col_action = true;
  p.has_change = true;
  return true;
}

// Single-doubleton path, doubleton_equations_scan=false (:1913-2084, :266-445, :1779-1909).
void dbl_scalar(Plan& p, const LpV& lp, const Params& prm) {
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0 || n == 0) return;
  const double zt = prm.zero_tol, tol = prm.bound_tol;
  const double rtol = std::fmax(1.0e-9, prm.bound_tol);
  std::vector<int32_t> nnz(n, 0);
  for (int32_t col = 0; col < n; ++col) {
    if (!p.keep_col[col]) continue;
    int32_t cnt = 0;
    for (int32_t q = lp.AT.ptr[col]; q < lp.AT.ptr[col + 1]; ++q)
      if (p.keep_row[lp.AT.idx[q]] && std::fabs(lp.AT.val[q]) > zt) ++cnt;
    nnz[col] = cnt;
  }
  const CsrV A = p.has_new_A ? p.new_A.view() : lp.A;
  int32_t target = m, ec = -1, kc = -1;
  double ev = 0.0, kv = 0.0;
  for (int32_t i = 0; i < m; ++i) {
    int32_t e, k2;
    double evv, kvv;
    if (!dbl_row_candidate(A, p, i, nnz, zt, tol, rtol, &e, &evv, &k2, &kvv)) continue;
    if (nnz[e] > nnz[k2] + prm.doubleton_equations_max_fill_in_proxy + 1) continue;
    if (dbl_fill_in(lp.AT, k2, e) > prm.doubleton_equations_max_fill_in_proxy) continue;
    target = i; ec = e; kc = k2; ev = evv; kv = kvv;
    break;  // atomicMin over rows selects the smallest row index
  }
  if (target >= m) return;
  const double rhs = p.AU[target];
  const double ol_e = p.l[ec], ou_e = p.u[ec], oc_e = p.c[ec], ol_k = p.l[kc], ou_k = p.u[kc];
  const double al = -kv / ev, be = rhs / ev;
  double ml = -INF, mu = INF;
  if (fin(ol_e)) { const double b = (ol_e - be) / al; if (al > 0.0) ml = b; else mu = b; }
  if (fin(ou_e)) { const double b = (ou_e - be) / al; if (al > 0.0) mu = b; else ml = b; }
  if (ml > mu) std::swap(ml, mu);
  const double nkl = fin(ml) ? std::fmax(ol_k, ml) : ol_k;
  const double nku = fin(mu) ? std::fmin(ou_k, mu) : ou_k;
  if (nkl > nku + tol) { p.has_infeasible = true; return; }
  for (int32_t t = 0; t < m; ++t) {
    double sh = 0.0;
    for (int32_t q = lp.AT.ptr[ec]; q < lp.AT.ptr[ec + 1]; ++q)
      if (lp.AT.idx[q] == t) { sh = lp.AT.val[q] * be; break; }
    p.AL[t] -= sh;
    p.AU[t] -= sh;
  }
  p.l[kc] = nkl;
  p.u[kc] = nku;
  p.c[kc] += oc_e * al;
  p.keep_row[target] = 0;
  p.keep_col[ec] = 0;
  if (prm.record_postsolve_tape)
    p.tape.append(DoubletonEquation, {ec, kc, target, 0},
                  {ev, kv, rhs, ol_e, ou_e, ol_k, ou_k, p.l[kc], p.u[kc], oc_e, 0.0, 0.0}, DualMinimal);
  // build_doubleton_scalar_new_A
  Csr out;
  out.rows = A.rows;
  out.cols = A.cols;
  out.ptr.assign(1, 0);
  for (int32_t r = 0; r < m; ++r) {
    const int32_t b = A.ptr[r], e = A.ptr[r + 1];
    if (r == target) {
      for (int32_t q = b; q < e; ++q) { out.idx.push_back(A.idx[q]); out.val.push_back(A.val[q]); }
      out.ptr.push_back(static_cast<int32_t>(out.idx.size()));
      continue;
    }
    bool ep = false, kp = false;
    double are = 0.0, okv = 0.0;
    for (int32_t q = b; q < e; ++q) {
      if (A.idx[q] == ec) { ep = true; are = A.val[q]; }
      else if (A.idx[q] == kc) { kp = true; okv = A.val[q]; }
    }
    const double nk = okv + al * are;
    const bool insert = ep && !kp && std::fabs(nk) > zt;
    bool inserted = !insert;
    for (int32_t q = b; q < e; ++q) {
      const int32_t col = A.idx[q];
      if (insert && !inserted && col > kc) { out.idx.push_back(kc); out.val.push_back(nk); inserted = true; }
      if (col == kc) {
        if (std::fabs(nk) > zt) { out.idx.push_back(kc); out.val.push_back(nk); }
      } else {
        out.idx.push_back(col);
        out.val.push_back(A.val[q]);
      }
    }
    if (insert && !inserted) { out.idx.push_back(kc); out.val.push_back(nk); }
    out.ptr.push_back(static_cast<int32_t>(out.idx.size()));
  }
  p.new_A = std::move(out);
  p.has_new_A = true;
  p.obj_delta += oc_e * be;
  p.has_row_action = true;
  p.has_col_action = true;
  p.has_change = true;
}
}  // namespace

// rules/rule_doubleton_equations.cu:1679-1777 (batch driver) and :1913-2084
void rule_doubleton_equations(Plan& p, const LpV& lp, const Stats&, const Params& prm) {
  if (p.has_infeasible || p.has_unbounded) return;
  if (!prm.doubleton_equations_scan) { dbl_scalar(p, lp, prm); return; }
  sort_rows_in_place(lp.AT);
  Csr at_own;
  CsrV at = lp.AT;
  const bool scan = prm.doubleton_equations_scan != 0;
  const int limit = scan && prm.doubleton_equations_max_batch_rounds > 0 &&
                            prm.doubleton_equations_max_batch_rounds < DBL_INNER_ROUNDS
                        ? prm.doubleton_equations_max_batch_rounds
                        : DBL_INNER_ROUNDS;
  std::vector<DblRec> recs;
  std::vector<double> objs;
  for (int round = 0; round < limit; ++round) {
    double os = 0.0, of = 0.0;
    const bool changed = dbl_batch_once(p, lp, at_own, at, prm, &recs, &os, &of);
    if (p.has_infeasible || p.has_unbounded) break;
    if (!changed) break;
    objs.push_back(os);
    objs.push_back(of);
    if (p.has_new_A) {
      at_own = transpose(p.new_A.view());
      at = at_own.view();
    }
  }
  if (prm.record_postsolve_tape)
    for (const DblRec& r : recs)
      p.tape.append(DoubletonEquation, {r.elim, r.keep, r.row, 0},
                    {r.ev, r.kv, r.au, r.el, r.eu, r.kl, r.ku, r.nkl, r.nku, r.ec, 0.0, 0.0}, DualMinimal);
  for (double o : objs) p.obj_delta += o;
}

// ---------------------------------------------------------------------------
// rules/rule_dual_fix.cu:174-554
void rule_dual_fix(Plan& p, const LpV& lp, const Stats&, const Params& prm, const uint8_t* cand) {
  if (p.has_infeasible || p.has_unbounded) return;
  const int32_t n = lp.A.cols, m = lp.A.rows;
  if (n == 0) return;
  const double zt = prm.zero_tol;
  std::vector<uint8_t> fixed(n, 0), inf_fix(n, 0);
  std::vector<double> v(n, 0.0), contrib(n, 0.0);
  int status = 0;  // atomicMax of 1 (infeasible) and 2 (unbounded)
  bool any = false, any_inf = false;
  for (int32_t j = 0; j < n; ++j) {
    if (!p.keep_col[j] || (cand && !cand[j])) continue;
    const double lj = p.l[j], uj = p.u[j];
    if (lj > uj + prm.bound_tol) { status = std::max(status, 1); continue; }
    const int32_t b = lp.AT.ptr[j], e = lp.AT.ptr[j + 1];
    if (std::fabs(p.c[j]) > zt && e - b > prm.dual_fix_max_nonzero_cost_col_degree) continue;
    bool down = false, up = false;
    for (int32_t q = b; q < e; ++q) {
      const int32_t row = lp.AT.idx[q];
      if (!p.keep_row[row]) continue;
      const double a = lp.AT.val[q];
      if (a > zt) { down = down || fin(p.AL[row]); up = up || fin(p.AU[row]); }
      else if (a < -zt) { down = down || fin(p.AU[row]); up = up || fin(p.AL[row]); }
      if (down && up) break;
    }
    const double cj = p.c[j];
    bool fx = false, infx = false;
    double vj = 0.0;
    if (cj > zt) {
      if (!down) { if (fin(lj)) { vj = lj; fx = true; } else { status = std::max(status, 2); continue; } }
    } else if (cj < -zt) {
      if (!up) { if (fin(uj)) { vj = uj; fx = true; } else { status = std::max(status, 2); continue; } }
    } else if (std::fabs(cj) <= zt) {
      if (!down) { fx = true; if (fin(lj)) vj = lj; else { vj = -INF; infx = true; } }
      else if (!up) { fx = true; if (fin(uj)) vj = uj; else { vj = INF; infx = true; } }
    }
    if (fx) {
      fixed[j] = 1;
      inf_fix[j] = infx ? 1 : 0;
      v[j] = vj;
      contrib[j] = infx ? 0.0 : cj * vj;
      any = true;
      any_inf = any_inf || infx;
    }
  }
  if (status == 1) { p.has_infeasible = true; return; }
  if (status == 2) { p.has_unbounded = true; return; }
  if (!any) return;
  std::vector<double> shift(m, 0.0);
  std::vector<uint8_t> row_remove(m, 0);
  bool row_changed = false;
  for (int32_t i = 0; i < m; ++i) {
    if (!p.keep_row[i]) continue;
    for (int32_t q = lp.A.ptr[i]; q < lp.A.ptr[i + 1]; ++q) {
      const int32_t j = lp.A.idx[q];
      if (fixed[j] && !inf_fix[j]) shift[i] += lp.A.val[q] * v[j];
      if (fixed[j] && inf_fix[j]) { row_remove[i] = 1; row_changed = true; }
    }
  }
  if (prm.record_postsolve_tape) {
    if (any_inf) {
      for (int32_t j = 0; j < n; ++j) {
        if (!fixed[j] || !inf_fix[j]) continue;
        const bool pos = v[j] > 0.0;
        std::vector<int32_t> idx{pos ? 1 : -1, j};
        std::vector<double> vals{0.0, pos ? p.l[j] : p.u[j]};
        int32_t active = 0;
        for (int32_t q = lp.AT.ptr[j]; q < lp.AT.ptr[j + 1]; ++q) {
          const int32_t row = lp.AT.idx[q];
          if (!p.keep_row[row]) continue;
          ++active;
          const int32_t rb = lp.A.ptr[row], re = lp.A.ptr[row + 1];
          idx.push_back(re - rb);
          vals.push_back(fin(p.AL[row]) ? p.AL[row] : p.AU[row]);
          for (int32_t t = rb; t < re; ++t) { idx.push_back(lp.A.idx[t]); vals.push_back(lp.A.val[t]); }
        }
        vals[0] = static_cast<double>(active);
        p.tape.append(FixedColInf, idx, vals, DualExact);
      }
    }
    if (prm.record_postsolve_tape_cpu) {
      std::vector<uint8_t> keep_after(m);
      for (int32_t i = 0; i < m; ++i) keep_after[i] = (p.keep_row[i] && !row_remove[i]) ? 1 : 0;
      fixed_col_tape(&p.tape, fixed.data(), inf_fix.data(), v.data(), keep_after.data(), p.c, lp.AT, n);
    }
  }
  for (int32_t j = 0; j < n; ++j) {
    if (!fixed[j]) continue;
    p.l[j] = v[j];
    p.u[j] = v[j];
    p.keep_col[j] = 0;
  }
  for (int32_t i = 0; i < m; ++i) {
    p.AL[i] -= shift[i];
    p.AU[i] -= shift[i];
    if (row_remove[i]) p.keep_row[i] = 0;
  }
  p.obj_delta += ordered_sum(contrib);
  p.has_row_action = p.has_row_action || row_changed;
  p.has_col_action = true;
  p.has_change = true;
}

// ---------------------------------------------------------------------------
// rules/rule_duplicate_columns.cu:95-842
void rule_duplicate_columns(Plan& p, const LpV& lp, const Stats&, const Params& prm) {
  if (p.has_infeasible || p.has_unbounded) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (n <= 1) return;
  if (lp.AT.nnz > 0) sort_rows_in_place(lp.AT);
  const CsrV& AT = lp.AT;
  const double zt = prm.zero_tol;
  const double ctol = std::fmax(prm.zero_tol, prm.bound_tol);
  const double otol = ctol;
  std::vector<uint64_t> h(n);
  for (int32_t j = 0; j < n; ++j) {
    if (!p.keep_col[j]) { h[j] = static_cast<uint64_t>(j); continue; }
    int32_t live = 0;
    double piv = 0.0;
    bool piv_set = false;
    for (int32_t q = AT.ptr[j]; q < AT.ptr[j + 1]; ++q) {
      if (p.keep_row[AT.idx[q]] && std::fabs(AT.val[q]) > zt) {
        ++live;
        if (!piv_set) { piv = AT.val[q]; piv_set = true; }
      }
    }
    if (!piv_set || live == 0) { h[j] = fnv_mix(FNV_OFFSET, static_cast<uint64_t>(j)); continue; }
    uint64_t x = fnv_mix(FNV_OFFSET, static_cast<uint64_t>(live));
    for (int32_t q = AT.ptr[j]; q < AT.ptr[j + 1]; ++q) {
      if (p.keep_row[AT.idx[q]] && std::fabs(AT.val[q]) > zt) {
        x = fnv_mix(x, static_cast<uint64_t>(static_cast<int64_t>(AT.idx[q])));
        x = fnv_mix(x, dbits(AT.val[q] / piv));
      }
    }
    h[j] = x;
  }
  std::vector<int32_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return h[a] != h[b] ? h[a] < h[b] : a < b; });
  // _duplicate_column_ratio (:181-229)
  auto next_live = [&](int32_t q, int32_t stop) {
    while (q < stop) {
      if (p.keep_row[AT.idx[q]] && std::fabs(AT.val[q]) > zt) return q;
      ++q;
    }
    return stop;
  };
  auto ratio_of = [&](int32_t j, int32_t k, double* out) -> bool {
    int32_t pj = AT.ptr[j], sj = AT.ptr[j + 1], pk = AT.ptr[k], sk = AT.ptr[k + 1];
    double r = 0.0;
    bool set = false;
    for (;;) {
      pj = next_live(pj, sj);
      pk = next_live(pk, sk);
      if (pj >= sj || pk >= sk) break;
      if (AT.idx[pj] != AT.idx[pk]) return false;
      const double aj = AT.val[pj], ak = AT.val[pk];
      if (!set) {
        if (std::fabs(aj) <= zt) return false;
        r = ak / aj;
        if (std::fabs(r) <= zt || !fin(r)) return false;
        set = true;
      }
      if (std::fabs(ak - r * aj) > ctol) return false;
      ++pj;
      ++pk;
    }
    pj = next_live(pj, sj);
    pk = next_live(pk, sk);
    if (pj < sj || pk < sk || !set) return false;
    *out = r;
    return true;
  };
  const double max_ratio = prm.duplicate_columns_max_abs_ratio;
  const bool use_guard = fin(max_ratio) && prm.duplicate_columns_scale_guard_min_groups > 1;
  auto ratio_violates = [](double amax, double r) {
    const double ar = std::fabs(r);
    return amax < 1.0 || ar > amax || ar * amax < 1.0;
  };
  std::vector<std::pair<int32_t, int32_t>> groups;
  for (int32_t s0 = 0; s0 < n;) {
    int32_t e0 = s0;
    while (e0 + 1 < n && h[order[e0 + 1]] == h[order[s0]]) ++e0;
    groups.emplace_back(s0, e0);
    s0 = e0 + 1;
  }
  int32_t guard_count = 0;
  if (use_guard) {
    // _kernel_count_small_high_ratio_parallel_groups (:284-344); only the
    // predicate count >= min_groups is observable.
    for (const auto& g : groups) {
      const int32_t sz = g.second - g.first + 1;
      if (sz < 2 || sz > prm.duplicate_columns_scale_guard_max_group_size) continue;
      bool hit = false;
      for (int32_t a = g.first; a < g.second && !hit; ++a) {
        const int32_t j = order[a];
        if (!p.keep_col[j]) continue;
        for (int32_t b = a + 1; b <= g.second; ++b) {
          const int32_t k = order[b];
          if (!p.keep_col[k]) continue;
          double r;
          if (!ratio_of(j, k, &r)) continue;
          if (std::fabs(p.c[k] - r * p.c[j]) > otol) continue;
          if (ratio_violates(max_ratio, r)) { hit = true; break; }
        }
      }
      if (hit) ++guard_count;
    }
  }
  const double active_max = (use_guard && guard_count < prm.duplicate_columns_scale_guard_min_groups)
                                ? INF : max_ratio;
  std::vector<uint8_t> del(n, 0), fixed(n, 0);
  std::vector<double> fixed_val(n, 0.0);
  struct Rec { int32_t del, kept; double r, dl, du, kl, ku; };
  std::vector<Rec> recs;
  bool unbounded = false, merged_any = false, fixed_any = false;
  // _kernel_duplicate_column_groups (:346-497)
  for (const auto& g : groups) {
    if (g.second <= g.first) continue;
    bool stop_group = false;
    for (int32_t a = g.first; a < g.second && !stop_group; ++a) {
      const int32_t j = order[a];
      if (!p.keep_col[j] || del[j] || fixed[j]) continue;
      for (int32_t b = a + 1; b <= g.second; ++b) {
        const int32_t k = order[b];
        if (!p.keep_col[k] || del[k] || fixed[k]) continue;
        double r;
        if (!ratio_of(j, k, &r)) continue;
        if (fin(active_max) && ratio_violates(active_max, r)) continue;
        const double gap = p.c[k] - r * p.c[j];
        const double tl = p.l[j], tu = p.u[j], sl = p.l[k], su = p.u[k];
        if (std::fabs(gap) <= otol) {
          recs.push_back(Rec{k, j, r, sl, su, tl, tu});
          double nl = -INF, nu = INF;
          if (r > 0.0) {
            if (fin(tl) && fin(sl)) nl = tl + r * sl;
            if (fin(tu) && fin(su)) nu = tu + r * su;
          } else {
            if (fin(tl) && fin(su)) nl = tl + r * su;
            if (fin(tu) && fin(sl)) nu = tu + r * sl;
          }
          p.l[j] = nl;
          p.u[j] = nu;
          del[k] = 1;
          merged_any = true;
          continue;
        }
        bool k_lo = false, k_hi = false, j_lo = false, j_hi = false;
        if (gap > otol) {
          if (r > 0.0) { k_lo = !fin(tu); j_hi = !fin(sl); }
          else { k_lo = !fin(tl); j_lo = !fin(sl); }
        } else {
          if (r > 0.0) { k_hi = !fin(tl); j_lo = !fin(su); }
          else { k_hi = !fin(tu); j_hi = !fin(su); }
        }
        if (k_lo) {
          if (!fin(sl)) { unbounded = true; stop_group = true; break; }
          fixed[k] = 1; fixed_val[k] = sl; fixed_any = true; continue;
        } else if (k_hi) {
          if (!fin(su)) { unbounded = true; stop_group = true; break; }
          fixed[k] = 1; fixed_val[k] = su; fixed_any = true; continue;
        }
        if (j_lo) {
          if (!fin(tl)) { unbounded = true; stop_group = true; break; }
          fixed[j] = 1; fixed_val[j] = tl; fixed_any = true; break;
        } else if (j_hi) {
          if (!fin(tu)) { unbounded = true; stop_group = true; break; }
          fixed[j] = 1; fixed_val[j] = tu; fixed_any = true; break;
        }
      }
    }
  }
  if (unbounded) { p.has_unbounded = true; return; }
  if (!(merged_any || fixed_any)) return;
  if (merged_any && prm.record_postsolve_tape) {
    std::sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) { return a.del < b.del; });
    for (const Rec& r : recs)
      p.tape.append(DuplicateColumn, {r.del, r.kept}, {r.r, r.dl, r.du, r.kl, r.ku}, DualMinimal);
  }
  std::vector<double> shift(m, 0.0);
  if (fixed_any) {
    if (prm.record_postsolve_tape)
      fixed_col_tape(&p.tape, fixed.data(), nullptr, fixed_val.data(), p.keep_row, p.c, lp.AT, n);
    for (int32_t i = 0; i < m; ++i) {
      if (!p.keep_row[i]) continue;
      for (int32_t q = lp.A.ptr[i]; q < lp.A.ptr[i + 1]; ++q)
        if (fixed[lp.A.idx[q]]) shift[i] += lp.A.val[q] * fixed_val[lp.A.idx[q]];
    }
    p.has_row_action = true;
  }
  std::vector<double> contrib;
  for (int32_t j = 0; j < n; ++j) {
    if (!(del[j] || fixed[j])) continue;
    p.keep_col[j] = 0;
    if (fixed[j]) {
      p.l[j] = fixed_val[j];
      p.u[j] = fixed_val[j];
      contrib.push_back(p.c[j] * fixed_val[j]);
    }
  }
  for (int32_t i = 0; i < m; ++i) { p.AL[i] -= shift[i]; p.AU[i] -= shift[i]; }
  p.obj_delta += ordered_sum(contrib);
  p.has_col_action = true;
  p.has_change = true;
}

// ---------------------------------------------------------------------------
// rules/rule_redundant_bounds.cu:31-411 (opt-in)
void rule_redundant_bounds(Plan& p, const LpV& lp, const Stats&, const Params& prm) {
  if (p.has_infeasible || p.has_unbounded) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0 || n == 0) return;
  const double zt = prm.zero_tol, tol = prm.bound_tol;
  const bool one_sided = prm.enable_redundant_bounds != 0;
  bool changed_any = false;
  for (;;) {
    std::vector<uint8_t> dlo(n, 0), dhi(n, 0), sl(n, 0), su(n, 0);
    // Rest activity of row i without column j, or false when a term is infinite.
    auto rest = [&](int32_t i, int32_t j, bool want_min, double* out) {
      double acc = 0.0;
      for (int32_t q = lp.A.ptr[i]; q < lp.A.ptr[i + 1]; ++q) {
        const int32_t col = lp.A.idx[q];
        if (col == j) continue;
        const double aq = lp.A.val[q];
        const double t = want_min ? (aq >= 0.0 ? aq * p.l[col] : aq * p.u[col])
                                  : (aq >= 0.0 ? aq * p.u[col] : aq * p.l[col]);
        if (!fin(t)) return false;
        acc += t;
      }
      *out = acc;
      return true;
    };
    for (int32_t j = 0; j < n; ++j) {
      const double lj = p.l[j], uj = p.u[j];
      const bool box = fin(lj) && fin(uj);
      const bool allow_box = box && (lp.c == nullptr || std::fabs(lp.c[j]) <= zt);
      const bool eligible = allow_box || (one_sided && !box);
      if (fin(uj) && eligible) {
        for (int32_t q = lp.AT.ptr[j]; q < lp.AT.ptr[j + 1]; ++q) {
          const int32_t i = lp.AT.idx[q];
          const double a = lp.AT.val[q];
          if (std::fabs(a) <= zt) continue;
          double iu = INF, r;
          if (a > zt && fin(p.AU[i])) { if (rest(i, j, true, &r)) iu = (p.AU[i] - r) / a; }
          else if (a < -zt && fin(p.AL[i])) { if (rest(i, j, false, &r)) iu = (p.AL[i] - r) / a; }
          if (fin(iu) && iu <= uj + tol) { dhi[j] = 1; break; }
        }
      }
      if (fin(lj) && eligible) {
        for (int32_t q = lp.AT.ptr[j]; q < lp.AT.ptr[j + 1]; ++q) {
          const int32_t i = lp.AT.idx[q];
          const double a = lp.AT.val[q];
          if (std::fabs(a) <= zt) continue;
          double il = -INF, r;
          if (a > zt && fin(p.AL[i])) { if (rest(i, j, false, &r)) il = (p.AL[i] - r) / a; }
          else if (a < -zt && fin(p.AU[i])) { if (rest(i, j, true, &r)) il = (p.AU[i] - r) / a; }
          if (fin(il) && il >= lj - tol) { dlo[j] = 1; break; }
        }
      }
    }
    std::vector<uint64_t> rkey(m, UINT64_MAX);
    for (int32_t i = 0; i < m; ++i) {
      for (int32_t q = lp.A.ptr[i]; q < lp.A.ptr[i + 1]; ++q) {
        const int32_t j = lp.A.idx[q];
        if (dlo[j] || dhi[j]) {
          const uint64_t key = (static_cast<uint64_t>(static_cast<uint32_t>(lp.AT.ptr[j + 1] - lp.AT.ptr[j])) << 32) |
                               static_cast<uint64_t>(static_cast<uint32_t>(j));
          rkey[i] = std::min(rkey[i], key);
        }
      }
    }
    bool any = false;
    for (int32_t j = 0; j < n; ++j) {
      if (!dlo[j] && !dhi[j]) continue;
      const uint64_t key = (static_cast<uint64_t>(static_cast<uint32_t>(lp.AT.ptr[j + 1] - lp.AT.ptr[j])) << 32) |
                           static_cast<uint64_t>(static_cast<uint32_t>(j));
      bool choose = true;
      for (int32_t q = lp.AT.ptr[j]; q < lp.AT.ptr[j + 1]; ++q)
        if (rkey[lp.AT.idx[q]] != key) { choose = false; break; }
      if (choose) { sl[j] = dlo[j]; su[j] = dhi[j]; any = true; }
    }
    if (!any) break;
    if (prm.record_postsolve_tape) {
      for (int32_t j = 0; j < n; ++j) {
        if (!sl[j] && !su[j]) continue;
        const double nl = sl[j] ? -INF : p.l[j];
        const double nu = su[j] ? INF : p.u[j];
        p.tape.append(BoundChangeNoRow, {j}, {p.l[j], p.u[j], nl, nu}, DualMinimal);
      }
    }
    for (int32_t j = 0; j < n; ++j) {
      if (sl[j]) p.l[j] = -INF;
      if (su[j]) p.u[j] = INF;
    }
    changed_any = true;
  }
  if (changed_any) { p.has_col_action = true; p.has_change = true; }
}

}  // namespace orc
