// CPU oracle: plan application, record maintenance, the tiered/fixed
// schedulers and postsolve replay -- TEST INFRASTRUCTURE.
// Restates cpp/src/presolve/gpu_presolve.cu and gpu_postsolve.cu of the
// reference (paths relative to PolyU-IOR/GPU-Presolver/).
#include "oracle_core.hpp"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <functional>
#include <stdexcept>

namespace orc {

LpV OwnedLp::view() const {
  LpV v;
  v.A = A.view();
  v.AT = AT.view();
  v.c = c.data();
  v.AL = AL.data();
  v.AU = AU.data();
  v.l = l.data();
  v.u = u.data();
  v.obj_constant = obj_constant;
  return v;
}

namespace {

inline bool fin(double x) { return std::isfinite(x); }

struct OwnedPlan {
  std::vector<uint8_t> keep_row, keep_col;
  std::vector<double> c, l, u, AL, AU;
  Plan p;
  explicit OwnedPlan(const OwnedLp& lp)
      : keep_row(lp.A.rows, 1), keep_col(lp.A.cols, 1), c(lp.c), l(lp.l), u(lp.u), AL(lp.AL), AU(lp.AU) {
    p.keep_row = keep_row.data();
    p.keep_col = keep_col.data();
    p.c = c.data();
    p.l = l.data();
    p.u = u.data();
    p.AL = AL.data();
    p.AU = AU.data();
  }
};

// recompute_{row,col}_stats_from_csr (gpu_presolve.cu:249-276, 696-732):
// counts include explicit zeros; the singleton entry is the first of the row.
struct OwnedStats {
  std::vector<int32_t> row_nnz, srow_col, col_nnz, scol_row;
  std::vector<uint8_t> srow_mask, empty_col, scol_mask;
  std::vector<double> srow_val, scol_val;
  Stats s;
  explicit OwnedStats(const OwnedLp& lp) {
    const int32_t m = lp.A.rows, n = lp.A.cols;
    row_nnz.resize(m); srow_col.resize(m); srow_mask.resize(m); srow_val.resize(m);
    for (int32_t i = 0; i < m; ++i) {
      const int32_t b = lp.A.ptr[i], cnt = lp.A.ptr[i + 1] - b;
      row_nnz[i] = cnt;
      srow_mask[i] = cnt == 1;
      srow_col[i] = cnt == 1 ? lp.A.idx[b] : -1;
      srow_val[i] = cnt == 1 ? lp.A.val[b] : 0.0;
    }
    col_nnz.resize(n); scol_row.resize(n); scol_mask.resize(n); scol_val.resize(n); empty_col.resize(n);
    for (int32_t j = 0; j < n; ++j) {
      const int32_t b = lp.AT.ptr[j], cnt = lp.AT.ptr[j + 1] - b;
      col_nnz[j] = cnt;
      empty_col[j] = cnt == 0;
      scol_mask[j] = cnt == 1;
      scol_row[j] = cnt == 1 ? lp.AT.idx[b] : -1;
      scol_val[j] = cnt == 1 ? lp.AT.val[b] : 0.0;
    }
    s.row_nnz = row_nnz.data();
    s.singleton_row_mask = srow_mask.data();
    s.singleton_row_col = srow_col.data();
    s.singleton_row_val = srow_val.data();
    s.col_nnz = col_nnz.data();
    s.empty_col_mask = empty_col.data();
    s.column_singleton_mask = scol_mask.data();
    s.column_singleton_row = scol_row.data();
    s.column_singleton_val = scol_val.data();
  }
};

int32_t map_idx(int32_t v, const std::vector<int32_t>& red2org) {
  if (v < 0 || v >= static_cast<int32_t>(red2org.size())) return v;
  return red2org[v];
}

// append_globalized_postsolve_tape (gpu_presolve.cu:1363-1533)
void append_global_tape(Tape& dst, const Tape& src, const std::vector<int32_t>& rmap,
                        const std::vector<int32_t>& cmap) {
  if (src.types.empty()) return;
  const int32_t base_rec = static_cast<int32_t>(dst.types.size());
  const int32_t base_idx = static_cast<int32_t>(dst.indices.size());
  const int32_t base_val = static_cast<int32_t>(dst.vals.size());
  (void)base_rec;
  dst.types.insert(dst.types.end(), src.types.begin(), src.types.end());
  dst.dual_modes.insert(dst.dual_modes.end(), src.dual_modes.begin(), src.dual_modes.end());
  dst.indices.insert(dst.indices.end(), src.indices.begin(), src.indices.end());
  dst.vals.insert(dst.vals.end(), src.vals.begin(), src.vals.end());
  for (size_t k = 1; k < src.index_starts.size(); ++k) {
    dst.index_starts.push_back(base_idx + src.index_starts[k]);
    dst.value_starts.push_back(base_val + src.value_starts[k]);
  }
  for (size_t k = 0; k < src.types.size(); ++k) {
    int32_t* ix = dst.indices.data() + base_idx + src.index_starts[k];
    const int32_t cnt = src.index_starts[k + 1] - src.index_starts[k];
    switch (src.types[k]) {
      case FixedCol:
        if (cnt > 0) {
          ix[0] = map_idx(ix[0], cmap);
          for (int32_t q = 1; q < cnt; ++q) ix[q] = map_idx(ix[q], rmap);
        }
        break;
      case FixedColInf:
        if (cnt >= 2) {
          ix[1] = map_idx(ix[1], cmap);
          int32_t q = 2;
          while (q < cnt) {
            const int32_t len = ix[q++];
            for (int32_t t = 0; t < len && q < cnt; ++t, ++q) ix[q] = map_idx(ix[q], cmap);
          }
        }
        break;
      case SubCol:
        if (cnt >= 3) {
          ix[0] = map_idx(ix[0], cmap);
          ix[1] = map_idx(ix[1], rmap);
          for (int32_t t = 0; t < ix[2] && 3 + t < cnt; ++t) ix[3 + t] = map_idx(ix[3 + t], cmap);
        }
        break;
      case DuplicateColumn:
        if (cnt >= 2) { ix[0] = map_idx(ix[0], cmap); ix[1] = map_idx(ix[1], cmap); }
        break;
      case DuplicateRow:
        if (cnt >= 2) { ix[0] = map_idx(ix[0], rmap); ix[1] = map_idx(ix[1], rmap); }
        break;
      case DeletedRow:
        if (cnt >= 1) {
          ix[0] = map_idx(ix[0], rmap);
          if (cnt >= 2) ix[1] = map_idx(ix[1], cmap);
        }
        break;
      case LhsChange:
      case RhsChange:
      case EqToIneq:
        if (cnt >= 1) ix[0] = map_idx(ix[0], rmap);
        break;
      case BoundChangeNoRow:
        if (cnt >= 1) ix[0] = map_idx(ix[0], cmap);
        break;
      case BoundChangeTheRow:
        if (cnt >= 2) { ix[0] = map_idx(ix[0], cmap); ix[1] = map_idx(ix[1], rmap); }
        break;
      case DoubletonEquation:
        if (cnt >= 4) {
          ix[0] = map_idx(ix[0], cmap);
          ix[1] = map_idx(ix[1], cmap);
          ix[2] = map_idx(ix[2], rmap);
          for (int32_t t = 0; t < ix[3] && 4 + t < cnt; ++t) ix[4 + t] = map_idx(ix[4 + t], rmap);
        }
        break;
      case FmeCol:
        if (cnt >= 3) {
          const int32_t inv = ix[0];
          ix[1] = map_idx(ix[1], cmap);
          for (int32_t t = 0; t < inv && 2 + t < cnt; ++t) ix[2 + t] = map_idx(ix[2 + t], rmap);
          const int32_t sp = 2 + inv;
          if (sp < cnt)
            for (int32_t t = 0; t < ix[sp] && sp + 1 + t < cnt; ++t) ix[sp + 1 + t] = map_idx(ix[sp + 1 + t], cmap);
        }
        break;
      default:
        break;
    }
  }
}

// update_record_from_plan (gpu_presolve.cu:1535-1802)
void update_record(Record& rec, const Plan& p, const OwnedLp& old, int32_t m_new, int32_t n_new,
                   double obj_new) {
  const int32_t m = old.A.rows, n = old.A.cols;
  std::vector<int32_t> next_rows, next_cols;
  if (m_new < m) {
    next_rows.reserve(m_new);
    for (int32_t i = 0; i < m; ++i) if (p.keep_row[i]) next_rows.push_back(rec.row_red2org[i]);
  }
  if (n_new < n) {
    next_cols.reserve(n_new);
    for (int32_t j = 0; j < n; ++j) {
      if (p.keep_col[j]) { next_cols.push_back(rec.col_red2org[j]); continue; }
      const double lo = p.l[j], hi = p.u[j];
      if (!fin(lo) || !fin(hi) || std::fabs(lo - hi) > 1.0e-9) continue;
      rec.fixed_idx.push_back(rec.col_red2org[j]);
      rec.fixed_val.push_back(0.5 * (lo + hi));
    }
  }
  append_global_tape(rec.tape, p.tape, rec.row_red2org, rec.col_red2org);
  if (m_new < m) rec.row_red2org = std::move(next_rows);
  if (n_new < n) rec.col_red2org = std::move(next_cols);
  rec.m1 = m_new;
  rec.n1 = n_new;
  rec.obj_constant_new = obj_new;
}

struct Working {
  OwnedLp lp;
  bool owns = false;
};

// apply_plan_to_working_lp (gpu_presolve.cu:2021-2147)
void apply_plan(Working& cur, const Plan& p, Record& rec) {
  const int32_t m = cur.lp.A.rows, n = cur.lp.A.cols;
  int32_t m_new = 0, n_new = 0;
  for (int32_t i = 0; i < m; ++i) m_new += p.keep_row[i] ? 1 : 0;
  for (int32_t j = 0; j < n; ++j) n_new += p.keep_col[j] ? 1 : 0;
  const double obj_new = cur.lp.obj_constant + p.obj_delta;
  const bool unchanged = cur.owns && !p.has_new_A && m_new == m && n_new == n;
  if (unchanged) {
    update_record(rec, p, cur.lp, m_new, n_new, obj_new);
    cur.lp.c.assign(p.c, p.c + n);
    cur.lp.l.assign(p.l, p.l + n);
    cur.lp.u.assign(p.u, p.u + n);
    cur.lp.AL.assign(p.AL, p.AL + m);
    cur.lp.AU.assign(p.AU, p.AU + m);
    cur.lp.obj_constant = obj_new;
    cur.owns = true;
    return;
  }
  const CsrV src = p.has_new_A ? p.new_A.view() : cur.lp.A.view();
  OwnedLp next;
  next.A = compact(src, p.keep_row, p.keep_col);
  next.A.cols = n_new;
  next.AT = transpose(next.A.view());
  for (int32_t j = 0; j < n; ++j) {
    if (!p.keep_col[j]) continue;
    next.c.push_back(p.c[j]);
    next.l.push_back(p.l[j]);
    next.u.push_back(p.u[j]);
  }
  for (int32_t i = 0; i < m; ++i) {
    if (!p.keep_row[i]) continue;
    next.AL.push_back(p.AL[i]);
    next.AU.push_back(p.AU[i]);
  }
  update_record(rec, p, cur.lp, m_new, n_new, obj_new);
  next.obj_constant = obj_new;
  cur.lp = std::move(next);
  cur.owns = true;
}

}  // namespace

namespace {
// kMinNnzForBatchedClosure (gpu_presolve.cu:2506); adjustable so the closure
// can be exercised on small LPs.
int32_t g_closure_min_nnz = 10000000;
// Times the closure ran since the process started (tests read it to show
// that a workload reaches the closure at the default threshold).
int64_t g_closure_runs = 0;
}  // namespace

void set_batched_closure_min_nnz(int32_t nnz) { g_closure_min_nnz = nnz; }
int32_t batched_closure_min_nnz() { return g_closure_min_nnz; }
int64_t batched_closure_runs() { return g_closure_runs; }

bool has_good_nnz_progress(int32_t before, int32_t after, double ratio) {
  if (before <= 0) return false;
  return static_cast<double>(after) < ratio * static_cast<double>(before);
}

// run_gpu_presolve_impl (gpu_presolve.cu:2162-3400) with the structure-specific
// rules absent (they find nothing on LPs without their block patterns).
Result presolve(const LpV& lp, const Params& prm, bool keep_reduced) {
  Result res;
  res.original_rows = lp.A.rows;
  res.original_cols = lp.A.cols;
  Record& rec = res.record;
  rec.m0 = rec.m1 = lp.A.rows;
  rec.n0 = rec.n1 = lp.A.cols;
  rec.row_red2org.resize(lp.A.rows);
  rec.col_red2org.resize(lp.A.cols);
  for (int32_t i = 0; i < lp.A.rows; ++i) rec.row_red2org[i] = i;
  for (int32_t j = 0; j < lp.A.cols; ++j) rec.col_red2org[j] = j;
  rec.obj_constant_old = rec.obj_constant_new = lp.obj_constant;

  Working cur;
  cur.lp.A = copy_csr(lp.A);
  cur.lp.AT = copy_csr(lp.AT);
  cur.lp.c.assign(lp.c, lp.c + lp.A.cols);
  cur.lp.l.assign(lp.l, lp.l + lp.A.cols);
  cur.lp.u.assign(lp.u, lp.u + lp.A.cols);
  cur.lp.AL.assign(lp.AL, lp.AL + lp.A.rows);
  cur.lp.AU.assign(lp.AU, lp.AU + lp.A.rows);
  cur.lp.obj_constant = lp.obj_constant;

  bool terminal = false;
  const auto t0 = std::chrono::steady_clock::now();
  const bool has_limit = prm.max_time > 0.0 && fin(prm.max_time);
  bool cs_dual_clean = false, cs_eq_clean = false;
  auto invalidate_clean = [&]() { cs_dual_clean = false; cs_eq_clean = false; };
  auto time_exceeded = [&]() {
    if (!has_limit) return false;
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >= prm.max_time;
  };
  auto finish_plan = [&](Plan& p) -> bool {
    res.has_infeasible = p.has_infeasible;
    res.has_unbounded = p.has_unbounded;
    terminal = p.has_infeasible || p.has_unbounded;
    const bool changed = !terminal && p.has_change;
    if (changed) {
      apply_plan(cur, p, rec);
      invalidate_clean();
    }
    return changed;
  };

  // run_row_phase (:2219-2301)
  auto row_phase = [&](bool empty, bool singleton, bool redundant, bool implied, bool dup) -> bool {
    if (terminal || time_exceeded()) return false;
    OwnedPlan op(cur.lp);
    OwnedStats os(cur.lp);
    const LpV v = cur.lp.view();
    Plan& p = op.p;
    if (!time_exceeded() && empty && prm.enable_empty_rows) rule_empty_rows(p, v, os.s, prm);
    if (!time_exceeded() && singleton && prm.enable_singleton_rows) rule_singleton_rows(p, v, os.s, prm);
    if (!time_exceeded() && redundant && prm.enable_infeasible_redundant_rows)
      rule_infeasible_redundant_rows(p, v, os.s, prm);
    if (!time_exceeded() && implied && prm.enable_implied_variable_bounds)
      rule_implied_variable_bounds(p, v, os.s, prm);
    if (!time_exceeded() && dup && prm.enable_duplicate_rows) rule_duplicate_rows(p, v, os.s, prm);
    return finish_plan(p);
  };

  // run_col_phase (:2303-2452)
  auto col_phase = [&](bool fixed, bool empty_cols, bool cs_dual, bool cs_eq, bool doubleton,
                       bool dual_fix, bool dup_cols, bool redundant_bounds) -> bool {
    if (terminal || time_exceeded()) return false;
    OwnedPlan op(cur.lp);
    OwnedStats os(cur.lp);
    const LpV v = cur.lp.view();
    Plan& p = op.p;
    const bool run_redundant =
        redundant_bounds && (prm.enable_redundant_bounds || prm.enable_zero_cost_redundant_box_bounds);
    if (!time_exceeded() && fixed && prm.enable_infeasible_fixed_variables) rule_infeasible_fixed_variables(p, v, prm);
    if (!time_exceeded() && empty_cols && prm.enable_empty_cols) rule_empty_cols(p, v, os.s, prm);
    if (!time_exceeded() && cs_dual && prm.enable_column_singletons_dual_infer)
      rule_column_singletons(p, v, os.s, prm, false, false);
    if (!time_exceeded() && cs_eq && prm.enable_column_singletons_eq)
      rule_column_singletons(p, v, os.s, prm, true, false);
    if (!time_exceeded() && doubleton && prm.enable_doubleton_equations) rule_doubleton_equations(p, v, os.s, prm);
    if (!time_exceeded() && dual_fix && prm.enable_dual_fix) rule_dual_fix(p, v, os.s, prm, nullptr);
    if (!time_exceeded() && dup_cols && prm.enable_duplicate_columns) rule_duplicate_columns(p, v, os.s, prm);
    if (!time_exceeded() && run_redundant) rule_redundant_bounds(p, v, os.s, prm);
    const bool changed = finish_plan(p);
    if (!changed) {
      const bool only_dual = !fixed && !empty_cols && cs_dual && !cs_eq && !doubleton && !dual_fix &&
                             !dup_cols && !redundant_bounds && prm.enable_column_singletons_dual_infer;
      const bool only_eq = !fixed && !empty_cols && !cs_dual && cs_eq && !doubleton && !dual_fix &&
                           !dup_cols && !redundant_bounds && prm.enable_column_singletons_eq;
      if (only_dual) cs_dual_clean = true;
      if (only_eq) cs_eq_clean = true;
    }
    return changed;
  };

  auto singleton_rows_to_exhaustion = [&]() {
    bool any = false;
    while (!terminal && !time_exceeded()) {
      const bool ch = row_phase(false, true, false, false, false);
      any = any || ch;
      if (!ch) break;
    }
    return any;
  };
  auto trivial_cleanup = [&]() {
    bool any = false;
    while (!terminal && !time_exceeded()) {
      bool pass = false;
      pass = col_phase(true, true, false, false, false, false, false, false) || pass;
      pass = row_phase(true, true, false, false, false) || pass;
      pass = col_phase(false, true, false, false, false, false, false, false) || pass;
      any = pass || any;
      if (!pass) break;
    }
    return any;
  };
  auto cleanup = [&]() {
    bool any = false;
    any = col_phase(true, true, false, false, false, false, false, false) || any;
    any = col_phase(false, false, false, false, false, true, false, false) || any;
    any = singleton_rows_to_exhaustion() || any;
    any = row_phase(true, false, false, false, false) || any;
    any = col_phase(false, true, false, false, false, false, false, false) || any;
    return any;
  };
  auto repeated_doubleton = [&]() {
    bool any = false;
    while (!terminal && !time_exceeded()) {
      const bool ch = col_phase(false, false, false, false, true, false, false, false);
      any = any || ch;
      if (!ch || prm.doubleton_equations_single_batch_per_iter) break;
    }
    return any;
  };
  // run_batched_singleton_cleanup_closure (:2503-2831): column-singleton EQ,
  // dual fix over incremental candidates and singleton rows, repeated on one
  // plan whose degree statistics are refreshed from the masks; a single
  // compaction at the end.
  auto singleton_closure = [&]() -> bool {
    if (terminal || time_exceeded() || cur.lp.A.nnz() < g_closure_min_nnz || !prm.enable_column_singletons_eq ||
        !prm.enable_dual_fix || !prm.enable_singleton_rows)
      return false;
    ++g_closure_runs;
    OwnedPlan op(cur.lp);
    OwnedStats os(cur.lp);
    const LpV v = cur.lp.view();
    Plan& p = op.p;
    const int32_t m = cur.lp.A.rows, n = cur.lp.A.cols;
    std::vector<uint8_t> row_before(op.keep_row), col_before(op.keep_col);
    std::vector<uint8_t> cand(n, 0), row_seen;
    std::vector<double> c_seen, l_seen, u_seen, AL_seen, AU_seen;
    auto snapshot = [&]() {
      row_seen = op.keep_row;
      c_seen = op.c;
      l_seen = op.l;
      u_seen = op.u;
      AL_seen = op.AL;
      AU_seen = op.AU;
    };
    snapshot();
    // _kernel_mark_changed_dual_cols, _kernel_mark_dual_cols_from_changed_rows (:200-247)
    auto mark_candidates = [&]() {
      for (int32_t j = 0; j < n; ++j)
        cand[j] = op.keep_col[j] && (op.c[j] != c_seen[j] || op.l[j] != l_seen[j] || op.u[j] != u_seen[j]);
      for (int32_t i = 0; i < m; ++i) {
        if (op.keep_row[i] == row_seen[i] && op.AL[i] == AL_seen[i] && op.AU[i] == AU_seen[i]) continue;
        for (int32_t q = v.A.ptr[i]; q < v.A.ptr[i + 1]; ++q)
          if (op.keep_col[v.A.idx[q]]) cand[v.A.idx[q]] = 1;
      }
    };
    // _kernel_update_{row,col}_degrees_from_removed_*, _kernel_refresh_* (:98-198)
    auto refresh = [&]() {
      for (int32_t j = 0; j < n; ++j) {
        if (!col_before[j] || op.keep_col[j]) continue;
        for (int32_t q = v.AT.ptr[j]; q < v.AT.ptr[j + 1]; ++q)
          if (op.keep_row[v.AT.idx[q]]) --os.row_nnz[v.AT.idx[q]];
      }
      for (int32_t i = 0; i < m; ++i) {
        if (!row_before[i] || op.keep_row[i]) continue;
        for (int32_t q = v.A.ptr[i]; q < v.A.ptr[i + 1]; ++q)
          if (op.keep_col[v.A.idx[q]]) --os.col_nnz[v.A.idx[q]];
      }
      row_before = op.keep_row;
      col_before = op.keep_col;
      for (int32_t i = 0; i < m; ++i) {
        const bool single = op.keep_row[i] && os.row_nnz[i] == 1;
        os.srow_mask[i] = single;
        os.srow_col[i] = -1;
        os.srow_val[i] = 0.0;
        if (!single) continue;
        for (int32_t q = v.A.ptr[i]; q < v.A.ptr[i + 1]; ++q) {
          if (!op.keep_col[v.A.idx[q]]) continue;
          os.srow_col[i] = v.A.idx[q];
          os.srow_val[i] = v.A.val[q];
          break;
        }
      }
      for (int32_t j = 0; j < n; ++j) {
        const bool single = op.keep_col[j] && os.col_nnz[j] == 1;
        os.scol_mask[j] = single;
        os.scol_row[j] = -1;
        os.scol_val[j] = 0.0;
        if (!single) continue;
        for (int32_t q = v.AT.ptr[j]; q < v.AT.ptr[j + 1]; ++q) {
          if (!op.keep_row[v.AT.idx[q]]) continue;
          os.scol_row[j] = v.AT.idx[q];
          os.scol_val[j] = v.AT.val[q];
          break;
        }
      }
    };
    bool any = false, first_dual = true;
    while (!terminal && !time_exceeded()) {
      p.has_change = false;
      rule_column_singletons(p, v, os.s, prm, true, true);
      const bool ch_eq = p.has_change;
      terminal = p.has_infeasible || p.has_unbounded;
      bool ch_dual = false;
      if (!terminal && !time_exceeded()) {
        if (first_dual) {
          std::fill(cand.begin(), cand.end(), uint8_t{1});
          first_dual = false;
        } else {
          mark_candidates();
        }
        snapshot();
        p.has_change = false;
        rule_dual_fix(p, v, os.s, prm, cand.data());
        ch_dual = p.has_change;
        terminal = p.has_infeasible || p.has_unbounded;
      }
      bool ch_rows = false;
      if (!terminal && !time_exceeded()) {
        refresh();
        p.has_change = false;
        rule_singleton_rows(p, v, os.s, prm);
        ch_rows = p.has_change;
        terminal = p.has_infeasible || p.has_unbounded;
        if (!terminal && ch_rows) refresh();
      }
      const bool ch = ch_eq || ch_dual || ch_rows;
      any = any || ch;
      if (!ch) break;
    }
    res.has_infeasible = p.has_infeasible;
    res.has_unbounded = p.has_unbounded;
    p.has_change = any;
    if (any && !terminal) {
      apply_plan(cur, p, rec);
      invalidate_clean();
    }
    return any;
  };
  // run_fast_phase (:2833-2863)
  auto fast_phase_fn = [&]() {
    bool any = singleton_closure();
    bool singleton = false;
    while (!terminal && !time_exceeded()) {
      const bool ch_dual = cs_dual_clean ? false : col_phase(false, false, true, false, false, false, false, false);
      const bool ch_eq = cs_eq_clean ? false : col_phase(false, false, false, true, false, false, false, false);
      const bool ch = ch_dual || ch_eq;
      any = any || ch;
      singleton = singleton || ch;
      if (ch) any = cleanup() || any;
      if (!ch) break;
    }
    const bool dbl = repeated_doubleton();
    any = dbl || any;
    if (dbl || singleton) any = cleanup() || any;
    return any;
  };
  // run_implied_variable_bounds_with_activity_closure (:2865-2987)
  auto implied_closure = [&]() {
    if (!prm.enable_implied_variable_bounds) return false;
    bool any = false;
    const int max_rounds = std::max(1, prm.implied_variable_bounds_max_rounds);
    const int cfg_bound_only = std::max(1, prm.implied_variable_bounds_max_bound_only_rounds);
    const int64_t pr = cur.lp.A.rows, pc = cur.lp.A.cols, pn = cur.lp.A.nnz();
    const bool wide_dense = pr >= 512 && pr <= 2048 && pc >= 128 * pr && pc <= 256 * pr &&
                            pn >= 2048 * pr && pn <= 8192 * pr;
    auto effective_limit = [&]() {
      int eff = cfg_bound_only;
      const int64_t budget = prm.implied_variable_bounds_bound_only_nnz_round_budget;
      if (budget <= 0 || cur.lp.A.nnz() <= 0) return eff;
      const int64_t br = std::max<int64_t>(1, budget / cur.lp.A.nnz());
      if (br < eff) eff = static_cast<int>(br);
      if (wide_dense) eff = std::max(eff, std::min(2, cfg_bound_only));
      return eff;
    };
    int rounds = 0, bound_only = 0, limit = cfg_bound_only;
    while (rounds < max_rounds && !terminal && !time_exceeded()) {
      limit = effective_limit();
      OwnedPlan op(cur.lp);
      OwnedStats os(cur.lp);
      const LpV v = cur.lp.view();
      Plan& p = op.p;
      bool group_changed = false, structural = false;
      while (rounds < max_rounds && !terminal && !time_exceeded()) {
        p.has_change = false;
        rule_implied_variable_bounds(p, v, os.s, prm);
        ++rounds;
        res.has_infeasible = p.has_infeasible;
        res.has_unbounded = p.has_unbounded;
        terminal = p.has_infeasible || p.has_unbounded;
        if (terminal || !p.has_change) break;
        group_changed = true;
        any = true;
        structural = p.has_new_A || p.has_row_action || p.has_col_action;
        if (structural) { bound_only = 0; break; }
        ++bound_only;
        if (bound_only >= limit) break;
      }
      if (group_changed && !terminal) {
        apply_plan(cur, p, rec);
        invalidate_clean();
      }
      if (terminal || !group_changed || !structural || bound_only >= limit) break;
    }
    if (any && !terminal && !time_exceeded()) any = row_phase(false, false, true, false, false) || any;
    return any;
  };
  auto medium_phase = [&]() {
    bool any = false, prop = false;
    prop = row_phase(false, false, true, false, false) || prop;
    prop = implied_closure() || prop;
    any = prop || any;
    if (prop) any = cleanup() || any;
    const bool dr = row_phase(false, false, false, false, true);
    any = dr || any;
    const bool dc = col_phase(false, false, false, false, false, false, true, false);
    any = dc || any;
    if (dr || dc) any = cleanup() || any;
    return any;
  };
  auto fixed_row_sequence = [&]() {
    bool any = false;
    any = row_phase(true, false, false, false, false) || any;
    any = row_phase(false, true, false, false, false) || any;
    any = row_phase(false, false, true, false, false) || any;
    const bool primal = implied_closure();
    any = primal || any;
    if (primal) any = trivial_cleanup() || any;
    any = row_phase(false, false, false, false, true) || any;
    return any;
  };
  auto trigger_col = [&](bool fixed, bool cs_dual, bool cs_eq, bool doubleton, bool dual_fix, bool dup) {
    const bool ch = col_phase(fixed, false, cs_dual, cs_eq, doubleton, dual_fix, dup, false);
    if (ch) (void)trivial_cleanup();
    return ch;
  };
  auto fixed_col_sequence = [&]() {
    bool any = false;
    any = trigger_col(true, false, false, false, false, false) || any;
    // structural L1 trigger (absent rule)
    any = col_phase(false, true, false, false, false, false, false, false) || any;
    bool ch = false;
    do { ch = trigger_col(false, true, false, false, false, false); any = ch || any; } while (ch && !terminal && !time_exceeded());
    do { ch = trigger_col(false, false, true, false, false, false); any = ch || any; } while (ch && !terminal && !time_exceeded());
    do {
      ch = trigger_col(false, false, false, true, false, false);
      any = ch || any;
    } while (ch && !prm.doubleton_equations_single_batch_per_iter && !terminal && !time_exceeded());
    any = trigger_col(false, false, false, false, true, false) || any;
    any = trigger_col(false, false, false, false, false, true) || any;
    return any;
  };

  if (!prm.use_tiered_scheduler) {
    for (int iter = 0; iter < prm.max_iters && !terminal && !time_exceeded(); ++iter) {
      bool ch = false;
      ch = fixed_row_sequence() || ch;
      ch = fixed_col_sequence() || ch;
      ++res.iterations;
      if (!ch) break;
    }
  } else {
    if (prm.enable_tiered_bootstrap && prm.max_iters > 0 && !time_exceeded()) {
      // Bootstrap (:3082-3332) without the structure-specific detectors.
      bool boot = false;
      boot = row_phase(true, false, false, false, false) || boot;
      boot = row_phase(false, true, false, false, false) || boot;
      boot = row_phase(false, false, true, false, false) || boot;
      boot = row_phase(false, false, false, false, true) || boot;
      bool ch = col_phase(true, true, false, false, false, false, false, false);
      boot = ch || boot;
      if (ch) boot = trivial_cleanup() || boot;
      boot = col_phase(false, true, false, false, false, false, false, false) || boot;
      do {
        ch = col_phase(false, false, true, false, false, false, false, false);
        boot = ch || boot;
        if (ch) boot = trivial_cleanup() || boot;
      } while (ch && !terminal && !time_exceeded());
      do {
        ch = col_phase(false, false, false, true, false, false, false, false);
        boot = ch || boot;
        if (ch) boot = trivial_cleanup() || boot;
      } while (ch && !terminal && !time_exceeded());
      ch = col_phase(false, false, false, false, false, true, false, false);
      boot = ch || boot;
      if (ch) boot = trivial_cleanup() || boot;
      ch = col_phase(false, false, false, false, false, false, true, false);
      boot = ch || boot;
      if (ch) boot = trivial_cleanup() || boot;
      (void)boot;
      ++res.iterations;
    }
    bool fast = true;
    int32_t cycle_nnz_before = cur.lp.A.nnz();
    constexpr double progress_ratio = 0.95;
    for (int iter = res.iterations; iter < prm.max_iters && !terminal && !time_exceeded(); ++iter) {
      bool ch = false;
      ch = cleanup() || ch;
      const int32_t nnz_before = cur.lp.A.nnz();
      if (fast) {
        ch = fast_phase_fn() || ch;
        fast = has_good_nnz_progress(nnz_before, cur.lp.A.nnz(), progress_ratio);
      } else {
        ch = medium_phase() || ch;
        const bool productive = has_good_nnz_progress(cycle_nnz_before, cur.lp.A.nnz(), progress_ratio);
        if (!ch || !productive) { ++res.iterations; break; }
        cycle_nnz_before = cur.lp.A.nnz();
        fast = true;
      }
      ++res.iterations;
      if (!ch && fast) break;
    }
  }
  if (!terminal && !time_exceeded() && (prm.enable_redundant_bounds || prm.enable_zero_cost_redundant_box_bounds))
    (void)col_phase(false, false, false, false, false, false, false, true);

  res.reduced_rows = cur.lp.A.rows;
  res.reduced_cols = cur.lp.A.cols;
  res.reduced_nnz = cur.lp.A.nnz();
  res.obj_constant_delta = cur.lp.obj_constant - lp.obj_constant;
  rec.m1 = res.reduced_rows;
  rec.n1 = res.reduced_cols;
  rec.obj_constant_new = cur.lp.obj_constant;
  res.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (keep_reduced) {
    res.reduced = std::move(cur.lp);
    res.has_reduced = true;
  }
  return res;
}

// ---------------------------------------------------------------------------
// Postsolve (gpu_postsolve.cu:143-1721)
namespace {

inline bool pfin(double x) { return x > -INF && x < INF; }
inline bool at_lower(double x, double l, double tol) { return pfin(l) && x <= l + tol; }
inline bool at_upper(double x, double u, double tol) { return pfin(u) && x >= u - tol; }
inline double clamp_inf(double x, double lo, double hi) {
  double y = x;
  if (pfin(lo)) y = std::fmax(y, lo);
  if (pfin(hi)) y = std::fmin(y, hi);
  return y;
}

// :189-210
double merged_source_value(double lo, double hi, double tol) {
  if (fin(lo) && fin(hi)) return 0.5 * (lo + hi);
  if (fin(lo)) {
    if (0.0 > lo + tol) return 0.0;
    return lo + std::fmax(1.0, std::fabs(lo) * std::fmax(tol, 1.0e-12));
  }
  if (fin(hi)) {
    if (0.0 < hi - tol) return 0.0;
    return hi - std::fmax(1.0, std::fabs(hi) * std::fmax(tol, 1.0e-12));
  }
  return 0.0;
}

// :240-282
void recover_dup_col(double y, double z_to, double r, double sl, double su, double tl, double tu, double tol,
                     double* x_from, double* x_to) {
  double lo = sl, hi = su;
  if (fin(tu)) { const double b = (y - tu) / r; if (r > 0.0) lo = std::fmax(lo, b); else hi = std::fmin(hi, b); }
  if (fin(tl)) { const double b = (y - tl) / r; if (r > 0.0) hi = std::fmin(hi, b); else lo = std::fmax(lo, b); }
  double lower = -INF, upper = INF;
  if (r > 0.0) {
    if (fin(tl) && fin(sl)) lower = tl + r * sl;
    if (fin(tu) && fin(su)) upper = tu + r * su;
  } else {
    if (fin(tl) && fin(su)) lower = tl + r * su;
    if (fin(tu) && fin(sl)) upper = tu + r * sl;
  }
  if (z_to > tol && fin(lower) && std::fabs(y - lower) <= tol) *x_from = r > 0.0 ? sl : su;
  else if (z_to < -tol && fin(upper) && std::fabs(y - upper) <= tol) *x_from = r > 0.0 ? su : sl;
  else *x_from = merged_source_value(lo, hi, tol);
  *x_to = y - r * (*x_from);
}

enum : uint8_t { ST_ZERO = 0, ST_NONNEG = 1, ST_NONPOS = 2, ST_FREE = 3 };
uint8_t row_state(double act, double lo, double hi, double tol) {
  const bool al = pfin(lo) && act <= lo + tol;
  const bool au = pfin(hi) && act >= hi - tol;
  if (al && au) return ST_FREE;
  if (al) return ST_NONNEG;
  if (au) return ST_NONPOS;
  return ST_ZERO;
}
void state_interval(uint8_t s, double* lo, double* hi) {
  if (s == ST_FREE) { *lo = -INF; *hi = INF; }
  else if (s == ST_NONNEG) { *lo = 0.0; *hi = INF; }
  else if (s == ST_NONPOS) { *lo = -INF; *hi = 0.0; }
  else { *lo = 0.0; *hi = 0.0; }
}
void deleted_dual_interval(double yp, double r, uint8_t ks, double* lo, double* hi) {
  const double piv = r * yp;
  if (ks == ST_FREE) { *lo = -INF; *hi = INF; }
  else if (ks == ST_ZERO) { *lo = piv; *hi = piv; }
  else if (ks == ST_NONNEG) { if (r > 0.0) { *lo = -INF; *hi = piv; } else { *lo = piv; *hi = INF; } }
  else { if (r > 0.0) { *lo = piv; *hi = INF; } else { *lo = -INF; *hi = piv; } }
}
// :409-469
bool doc_dual_split(double yp, double r, double kal, double kau, double dsl, double dsu, double tol, double* op,
                    double* od) {
  if (std::fabs(yp) <= tol) { *op = 0.0; *od = 0.0; return true; }
  const double bl = std::fmax(kal, dsl), bu = std::fmin(kau, dsu);
  if (!(bl <= bu + tol)) { *op = yp; *od = 0.0; return false; }
  const bool lt = bl > kal + tol, ut = bu < kau - tol;
  if (lt && ut) { *op = yp; *od = 0.0; return false; }
  if (!lt && !ut) { *op = yp; *od = 0.0; return true; }
  if (yp > 0.0) {
    if (lt) { *op = 0.0; *od = r * yp; } else { *op = yp; *od = 0.0; }
    return true;
  }
  if (lt) { *op = yp; *od = 0.0; } else { *op = 0.0; *od = r * yp; }
  return true;
}
// :471-515
void dbl_recover_z(double xk, double zk_hat, double al, double okl, double oku, double nkl, double nku, double ls,
                   double us, double tol, double* zk, double* ze) {
  if (std::fabs(al) <= tol) { *zk = zk_hat; *ze = 0.0; return; }
  const bool atl = pfin(nkl) && xk <= nkl + tol;
  const bool atu = pfin(nku) && xk >= nku - tol;
  if (!atl && !atu) { *zk = 0.0; *ze = 0.0; return; }
  double src = atl ? ls : us;
  if (src == 0.0) {
    if (atl) src = (pfin(okl) && std::fabs(nkl - okl) <= tol) ? 1.0 : 2.0;
    else src = (pfin(oku) && std::fabs(nku - oku) <= tol) ? 1.0 : 2.0;
  }
  if (src == 2.0) { *zk = 0.0; *ze = zk_hat / al; }
  else { *zk = zk_hat; *ze = 0.0; }
}

// _kernel_replay_postsolve_tape (:605-1046), one record at a time in reverse.
void replay(const RecordV& R, double* x, double* y, double* z, uint8_t* zret, const LpV* orig, int32_t begin,
            int32_t end, bool primal, bool dual, double tol) {
  const bool has_orig = orig != nullptr;
  for (int32_t k = end - 1; k >= begin; --k) {
    const int32_t type = R.types[k];
    const int32_t i0 = R.index_starts[k], i1 = R.index_starts[k + 1] - 1;
    const int32_t v0 = R.value_starts[k], v1 = R.value_starts[k + 1] - 1;
    const int32_t* I = R.indices;
    const double* V = R.vals;
    if (type == DuplicateColumn) {
      const int32_t from = I[i0], to = I[i0 + 1];
      const double r = V[v0];
      if (primal) {
        double xf = 0.0, xt = x[to];
        recover_dup_col(x[to], z ? z[to] : 0.0, r, V[v0 + 1], V[v0 + 2], V[v0 + 3], V[v0 + 4], tol, &xf, &xt);
        x[from] = xf;
        x[to] = xt;
      }
      if (dual && z) { z[from] = r * z[to]; if (zret) zret[from] = 1; }
    } else if (type == DoubletonEquation) {
      const int32_t ec = I[i0], kc = I[i0 + 1], row = I[i0 + 2];
      const double ecoef = V[v0], kcoef = V[v0 + 1], rhs = V[v0 + 2], eobj = V[v0 + 9];
      const double al = -kcoef / ecoef, be = rhs / ecoef;
      if (primal) x[ec] = be + al * x[kc];
      if (dual && z && y && has_orig) {
        const double ls = v1 >= v0 + 10 ? V[v0 + 10] : 0.0;
        const double us = v1 >= v0 + 11 ? V[v0 + 11] : 0.0;
        double zk = 0.0, ze = 0.0;
        dbl_recover_z(x[kc], z[kc], al, V[v0 + 5], V[v0 + 6], V[v0 + 7], V[v0 + 8], ls, us, tol, &zk, &ze);
        const bool payload = i1 >= i0 + 3 && v1 >= v0 + 11;
        double yrow;
        if (payload) {
          double so = 0.0;
          for (int32_t t = 0; t < I[i0 + 3]; ++t) so += V[v0 + 12 + t] * y[I[i0 + 4 + t]];
          yrow = (eobj - so - ze) / ecoef;
        } else {
          double so = 0.0;
          for (int32_t q = orig->AT.ptr[ec]; q < orig->AT.ptr[ec + 1]; ++q)
            if (orig->AT.idx[q] != row) so += orig->AT.val[q] * y[orig->AT.idx[q]];
          yrow = (orig->c[ec] - so - ze) / ecoef;
        }
        y[row] = yrow;
        z[kc] = zk;
        z[ec] = ze;
        if (zret) { zret[kc] = 1; zret[ec] = 1; }
      } else if (dual) {
        if (z) { z[ec] = 0.0; if (zret) zret[ec] = 1; }
        if (y && row >= 0) y[row] = eobj / ecoef;
      }
    } else if (type == SubCol) {
      const int32_t ec = I[i0], row = I[i0 + 1], sc = I[i0 + 2];
      const double piv = V[v0], rhs = V[v0 + 1], eobj = V[v0 + 4];
      const bool row_deleted = V[v0 + 5] != 0.0;
      double part = 0.0;
      for (int32_t t = 0; t < sc; ++t) part += V[v0 + 6 + t] * x[I[i0 + 3 + t]];
      if (primal) x[ec] = (rhs - part) / piv;
      if (dual && y && row >= 0) {
        if (row_deleted) y[row] = eobj / piv; else y[row] += eobj / piv;
      }
      if (dual && z) {
        z[ec] = (row_deleted || !y) ? 0.0 : eobj - piv * y[row];
        if (zret) zret[ec] = 1;
      }
    } else if (type == FixedCol) {
      const int32_t col = I[i0];
      if (primal) x[col] = V[v0];
      if (dual && z) {
        double zj = V[v0 + 1];
        for (int32_t pos = i0 + 1; pos <= i1; ++pos) {
          const double coef = V[v0 + 2 + (pos - (i0 + 1))];
          if (y) zj -= coef * y[I[pos]];
        }
        z[col] = zj;
        if (zret) zret[col] = 1;
      }
    } else if (type == FixedColInf) {
      const bool pos_inf = I[i0] > 0;
      const int32_t col = I[i0 + 1];
      const int32_t nrows = static_cast<int32_t>(std::llround(V[v0]));
      double ext = V[v0 + 1];
      int32_t ip = i0 + 2, vp = v0 + 2;
      if (primal) {
        for (int32_t r = 0; r < nrows; ++r) {
          const int32_t len = I[ip];
          double side = V[vp], coef = 0.0;
          for (int32_t t = 0; t < len; ++t) {
            const int32_t rc = I[ip + 1 + t];
            const double rv = V[vp + 1 + t];
            if (rc == col) coef = rv; else side -= rv * x[rc];
          }
          const double cand = side / coef;
          ext = pos_inf ? std::fmax(ext, cand) : std::fmin(ext, cand);
          ip += len + 1;
          vp += len + 1;
        }
        x[col] = ext;
      }
      if (dual && z) { z[col] = 0.0; if (zret) zret[col] = 1; }
    } else if (type == FmeCol) {
      const int32_t K = I[i0], ec = I[i0 + 1];
      const int32_t L = static_cast<int32_t>(std::llround(V[v0]));
      int32_t cc = i0 + 3 + K, vp = v0 + 1;
      double lo = -INF, hi = INF;
      bool hl = false, hh = false;
      if (primal) {
        for (int32_t s = 0; s < L; ++s) {
          const double orient = V[vp];
          const int32_t len = static_cast<int32_t>(std::llround(V[vp + 1]));
          double cand = V[vp + 2];
          for (int32_t t = 0; t < len; ++t) cand += V[vp + 3 + t] * x[I[cc++]];
          if (orient > 0.0) { lo = std::fmax(lo, cand); hl = true; }
          else { hi = std::fmin(hi, cand); hh = true; }
          vp += 3 + len;
        }
        x[ec] = hl ? lo : (hh ? hi : 0.0);
      }
      if (dual && z) { z[ec] = 0.0; if (zret) zret[ec] = 1; }
      if (dual && y) for (int32_t t = 0; t < K; ++t) y[I[i0 + 2 + t]] = 0.0;
    } else if (dual && type == EqToIneq) {
      if (y) y[I[i0]] += V[v0];
    } else if (dual && type == DuplicateRow && i1 >= i0 + 1 && v1 >= v0 + 4) {
      const int32_t kr = I[i0], dr = I[i0 + 1];
      const double r = V[v0], kal = V[v0 + 1], kau = V[v0 + 2], dal = V[v0 + 3], dau = V[v0 + 4];
      if (std::fabs(r) <= tol || !y) continue;
      const double a = dal * r, b = dau * r;
      const double dsl = r >= 0.0 ? a : b, dsu = r >= 0.0 ? b : a;
      if (dsl <= kal + tol && kau <= dsu + tol) { y[dr] = 0.0; continue; }
      if (has_orig) {
        double ap = 0.0, ad = 0.0;
        for (int32_t q = orig->A.ptr[kr]; q < orig->A.ptr[kr + 1]; ++q) ap += orig->A.val[q] * x[orig->A.idx[q]];
        for (int32_t q = orig->A.ptr[dr]; q < orig->A.ptr[dr + 1]; ++q) ad += orig->A.val[q] * x[orig->A.idx[q]];
        const uint8_t ks = row_state(ap, kal, kau, tol), ds = row_state(ad, dal, dau, tol);
        double ydl, ydh, ypl, yph;
        state_interval(ds, &ydl, &ydh);
        deleted_dual_interval(y[kr], r, ks, &ypl, &yph);
        const double sl = std::fmax(ydl, ypl), sh = std::fmin(ydh, yph);
        if (sl <= sh + tol) {
          const double ydn = clamp_inf(0.0, sl, sh);
          const double ypn = y[kr] - ydn / r;
          y[kr] = ypn;
          y[dr] = ydn;
          continue;
        }
        double ypd = y[kr], ydd = 0.0;
        if (doc_dual_split(y[kr], r, kal, kau, dsl, dsu, tol, &ypd, &ydd)) { y[kr] = ypd; y[dr] = ydd; }
      }
    } else if (dual && type == DeletedRow && has_orig && i1 >= i0 + 1 && v1 >= v0 + 2) {
      const int32_t row = I[i0], col = I[i0 + 1];
      const double oal = V[v0], oau = V[v0 + 1], aij = V[v0 + 2];
      if (std::fabs(aij) <= tol) continue;
      double gamma = orig->c[col];
      for (int32_t q = orig->AT.ptr[col]; q < orig->AT.ptr[col + 1]; ++q)
        if (orig->AT.idx[q] != row && y) gamma -= orig->AT.val[q] * y[orig->AT.idx[q]];
      const double xj = x[col], zbar = z[col];
      const double act = aij * xj;
      const bool ral = pfin(oal) && act <= oal + tol;
      const bool rau = pfin(oau) && act >= oau - tol;
      double izl = -INF, izh = INF;
      if (!(ral && rau)) {
        if (ral) { if (aij > 0.0) izh = gamma; else izl = gamma; }
        else if (rau) { if (aij > 0.0) izl = gamma; else izh = gamma; }
        else { izl = gamma; izh = gamma; }
      }
      const bool atl = at_lower(xj, orig->l[col], tol), atu = at_upper(xj, orig->u[col], tol);
      double zl = izl, zh = izh;
      if (!atl && !atu) {
        if ((pfin(izl) && izl > tol) || (pfin(izh) && izh < -tol)) { zl = gamma; zh = gamma; }
        else { zl = 0.0; zh = 0.0; }
      } else if (atl && !atu) zl = std::fmax(zl, 0.0);
      else if (!atl && atu) zh = std::fmin(zh, 0.0);
      if (pfin(zl) && pfin(zh) && zl > zh + tol) continue;
      const double zn = clamp_inf(zbar, zl, zh);
      z[col] = zn;
      if (y) y[row] = (gamma - zn) / aij;
      if (zret) zret[col] = 1;
    } else if (dual && type == BoundChangeTheRow && has_orig && i1 >= i0 + 1) {
      const int32_t col = I[i0], row = I[i0 + 1];
      if (zret && !zret[col]) continue;
      const double ol = V[v0], ou = V[v0 + 1], nl = V[v0 + 2], nu = V[v0 + 3];
      const bool lc = nl > ol + tol, uc = nu < ou - tol;
      if (lc == uc) continue;
      const double implied = lc ? nl : nu;
      const double other = lc ? ou : ol;
      const bool other_is_lower = !lc;
      const double xj = x[col], zj = z[col];
      if (pfin(other) && std::fabs(xj - other) <= tol) {
        const bool ok = other_is_lower ? (zj >= -tol) : (zj <= tol);
        if (ok) continue;
      }
      if (std::fabs(xj - implied) > tol) continue;
      double aij = 0.0;
      for (int32_t q = orig->A.ptr[row]; q < orig->A.ptr[row + 1]; ++q)
        if (orig->A.idx[q] == col) { aij = orig->A.val[q]; break; }
      if (std::fabs(aij) <= tol) continue;
      if (y) y[row] += zj / aij;
      for (int32_t q = orig->A.ptr[row]; q < orig->A.ptr[row + 1]; ++q) {
        const int32_t kc = orig->A.idx[q];
        if (kc == col || (zret && !zret[kc])) continue;
        z[kc] -= (orig->A.val[q] / aij) * zj;
      }
      z[col] = 0.0;
      if (zret) zret[col] = 1;
    }
  }
}

}  // namespace

// postsolve_gpu (gpu_postsolve.cu:1638-1721) for records without
// structure-rule checkpoints.
void postsolve(const RecordV& R, const double* x_red, const double* y_red, const double* z_red, const LpV* orig,
               double* x, double* y, double* z) {
  const double tol = 1.0e-7;
  const int32_t n0 = R.n0, m0 = R.m0;
  std::fill(x, x + n0, 0.0);
  std::fill(z, z + n0, 0.0);
  if (y) std::fill(y, y + m0, 0.0);
  std::vector<uint8_t> zret(n0, 0);
  for (int32_t j = 0; j < R.n1; ++j) {
    const int32_t o = R.col_red2org[j];
    if (x_red) x[o] = x_red[j];
    if (z_red) z[o] = z_red[j];
    zret[o] = 1;
  }
  if (y_red && y)
    for (int32_t i = 0; i < R.m1; ++i) y[R.row_red2org[i]] = y_red[i];
  for (int32_t k = 0; k < R.fixed_count; ++k) x[R.fixed_idx[k]] = R.fixed_val[k];
  double* yy = m0 > 0 ? y : nullptr;
  replay(R, x, yy, z, zret.data(), orig, 0, R.tape_count, true, false, tol);
  replay(R, x, yy, z, zret.data(), orig, 0, R.tape_count, false, true, tol);
  // project_column_duals_from_original (:1370-1435)
  if (!orig || !yy || orig->AT.rows <= 0 || !orig->c || !orig->l || !orig->u) return;
  std::vector<uint8_t> prot(n0, 0);
  for (int32_t k = 0; k < R.tape_count; ++k) {
    const int32_t i0 = R.index_starts[k];
    const int32_t t = R.types[k];
    auto mark = [&](int32_t c) { if (0 <= c && c < n0) prot[c] = 1; };
    if (t == SubCol) mark(R.indices[i0]);
    else if (t == DoubletonEquation) { mark(R.indices[i0]); mark(R.indices[i0 + 1]); }
    else if (t == FixedColInf) mark(R.indices[i0 + 1]);
    else if (t == DuplicateColumn) mark(R.indices[i0]);
  }
  for (int32_t col = 0; col < orig->AT.rows; ++col) {
    if (prot[col]) continue;
    double aty = 0.0;
    for (int32_t q = orig->AT.ptr[col]; q < orig->AT.ptr[col + 1]; ++q) aty += orig->AT.val[q] * yy[orig->AT.idx[q]];
    const double r = orig->c[col] - aty;
    const double xj = x[col], lj = orig->l[col], uj = orig->u[col];
    const bool atl = at_lower(xj, lj, tol), atu = at_upper(xj, uj, tol);
    double zj;
    if (atl && atu) zj = r;
    else if (!atl && !atu) zj = 0.0;
    else if (atl) zj = std::fmax(r, 0.0);
    else zj = std::fmin(r, 0.0);
    const double eps = std::fmin(tol, 1.0e-12);
    if (zj > eps && !fin(lj)) zj = 0.0;
    else if (zj < -eps && !fin(uj)) zj = 0.0;
    z[col] = zj;
  }
}

}  // namespace orc
