// Rule operator API (reference cpp/include/gpu_presolver/presolve/rules/*.hpp):
// host orchestration of each reduction rule over the HIP kernels of
// k_rules_rows.hip / k_rules_cols.hip.  Each apply_rule_* reads the plan and
// the phase-entry LP exactly as the reference rule of the same name does and
// leaves the same plan state, flags and tape records.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <vector>

#include "gp_host.hpp"
#include "gpu_presolver/presolve/fixed_col_tape.hpp"
#include "gpu_presolver/presolve/gpu_presolve_kernels.hpp"
#include "gpu_presolver/presolve/rules/rule_column_singletons.hpp"
#include "gpu_presolver/presolve/rules/rule_doubleton_equations.hpp"
#include "gpu_presolver/presolve/rules/rule_dual_fix.hpp"
#include "gpu_presolver/presolve/rules/rule_duplicate_columns.hpp"
#include "gpu_presolver/presolve/rules/rule_duplicate_rows.hpp"
#include "gpu_presolver/presolve/rules/rule_empty_cols.hpp"
#include "gpu_presolver/presolve/rules/rule_empty_rows.hpp"
#include "gpu_presolver/presolve/rules/rule_implied_variable_bounds.hpp"
#include "gpu_presolver/presolve/rules/rule_infeasible_fixed_variables.hpp"
#include "gpu_presolver/presolve/rules/rule_infeasible_redundant_rows.hpp"
#include "gpu_presolver/presolve/rules/rule_redundant_bounds.hpp"
#include "gpu_presolver/presolve/rules/rule_singleton_rows.hpp"

namespace gpu_presolver::presolve {

using amd::check;
using amd::ctx;
using amd::DBuf;
using amd::csr_view;
using amd::plan_view;
using amd::row_activity;
using amd::append_record;
using RT = PostsolveReductionType;
using DM = PostsolveDualMode;

namespace {

template <class T>
void d2d(T* dst, const T* src, int64_t n, const char* context) {
  if (n > 0)
    check(hipMemcpyAsync(dst, src, sizeof(T) * static_cast<size_t>(n), hipMemcpyDeviceToDevice, ctx().stream()),
          context);
}

template <class T>
void zero(T* dst, int64_t n, const char* context) {
  amd::CopyBatch().zero(dst, n).run(context);
}

// Packs the indices i with mask[i] != 0 (ascending) and copies them to the host.
// Ascending indices of the set mask entries; the first 4096 travel with
// the count, so a short list costs one wait.
std::vector<int32_t> masked_indices(const uint8_t* mask, int32_t n) {
  if (n <= 0) return {};
  auto& c = ctx();
  hipStream_t s = c.stream();
  DBuf<int32_t> prefix(static_cast<int64_t>(n) + 1), out(n);
  check(gpk_scan_lb_mask(mask, prefix, n, c.lookback(gpk_lookback_tiles(n)), s), "scan mask");
  check(gpk_pack_index(mask, prefix, n, out, s), "pack indices");
  const int32_t head = std::min<int32_t>(n, 4096);
  std::vector<int32_t> h(static_cast<size_t>(head));
  c.download(h.data(), out.get(), sizeof(int32_t) * static_cast<size_t>(head), "indices to host");
  const int32_t k = c.scalar_to_host(prefix.get() + n, "index count");
  if (k > head) {
    h.resize(static_cast<size_t>(k));
    c.download(h.data() + head, out.get() + head, sizeof(int32_t) * static_cast<size_t>(k - head), "indices to host");
    c.sync("indices to host");
  }
  h.resize(static_cast<size_t>(k));
  return h;
}

}  // namespace

// gpu_presolve_kernels.cu:86-111
void compute_row_activity_summary(double* row_min_fin, double* row_max_fin, std::int32_t* row_min_neg_inf_count,
                                  std::int32_t* row_max_pos_inf_count, const DeviceCsrMatrix& A_csr, const double* l,
                                  const double* u, double zero_tol) {
  amd::AbandonOnThrow abandon_guard;
  if (A_csr.rows == 0) return;
  row_activity(A_csr.rowPtr, A_csr.colVal, A_csr.nzVal, A_csr.rows, A_csr.cols, l, u, zero_tol, row_min_fin,
               row_max_fin, row_min_neg_inf_count, row_max_pos_inf_count, ctx().stream());
}

// gpu_presolve_kernels.cu:113-124
void compute_col_max_abs(double* col_max_abs, const DeviceCsrMatrix& AT_csr) {
  amd::AbandonOnThrow abandon_guard;
  if (AT_csr.rows == 0) return;
  check(gpk_col_max_abs(AT_csr.rowPtr, AT_csr.nzVal, AT_csr.rows, col_max_abs, ctx().stream()), "col max abs");
}

// fixed_col_tape.cu:107-267
double append_compact_fixed_col_tape_from_device(PostsolveTape* tape, const std::uint8_t* fixed_mask,
                                                 const std::uint8_t* excluded_mask, const double* fixed_val,
                                                 const std::uint8_t* keep_row, const double* c,
                                                 const DeviceCsrMatrix& AT, std::int32_t n, const char* context) {
  amd::AbandonOnThrow abandon_guard;
  (void)context;
  return amd::fixed_col_tape(tape, fixed_mask, excluded_mask, fixed_val, keep_row, c, AT, n);
}

// ---------------------------------------------------------------------------
// rule_empty_rows.cu:188-240
void apply_rule_empty_rows(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveStatsGpu& stats,
                           const PresolveParams& prm) {
  amd::AbandonOnThrow abandon_guard;
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t m = lp.A.rows;
  if (m == 0) return;
  auto& c = ctx();
  int32_t* f = c.flags();
  check(gpk_empty_rows(m, plan_view(plan), stats.row_nnz, prm.feasibility_tol, 0, f, c.stream()), "empty rows");
  const int32_t* h = c.read_flags(2);
  if (h[0]) {
    plan.has_infeasible = true;
    return;
  }
  if (h[1]) {
    check(gpk_empty_rows(m, plan_view(plan), stats.row_nnz, prm.feasibility_tol, 1, f, c.stream()), "empty rows apply");
    plan.has_row_action = true;
    plan.has_change = true;
  }
}

// rule_singleton_rows.cu:127-248
void apply_rule_singleton_rows(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveStatsGpu& stats,
                               const PresolveParams& prm) {
  amd::AbandonOnThrow abandon_guard;
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0) return;
  auto& c = ctx();
  hipStream_t s = c.stream();
  DBuf<double> cl(n), cu(n);
  DBuf<uint8_t> remove(m);
  amd::CopyBatch().copy(cl.get(), plan.new_l, n).copy(cu.get(), plan.new_u, n).run("singleton rows candidates");
  const gpk_plan P = plan_view(plan);
  check(gpk_singleton_rows_bounds(m, n, P, stats.singleton_row_mask, stats.singleton_row_col, stats.singleton_row_val,
                                  prm.zero_tol, cl, cu, remove, s),
        "singleton rows bounds");
  int32_t* f = c.flags();
  check(gpk_singleton_rows_check(m, n, P, cl, cu, remove, prm.bound_tol, f, s), "singleton rows check");
  const int32_t* h = c.read_flags(3);
  if (h[0]) {
    plan.has_infeasible = true;
    return;
  }
  const bool removed = h[1] != 0, changed = h[2] != 0;
  if (!(removed || changed)) return;
  if (prm.record_postsolve_tape && prm.record_postsolve_tape_cpu && removed) {
    std::vector<uint8_t> hr;
    std::vector<int32_t> hc;
    std::vector<double> hv, hal, hau;
    c.begin_copy_group();
    c.to_host_async(hr, remove.get(), m, "singleton tape rows");
    c.to_host_async(hc, stats.singleton_row_col, m, "singleton tape cols");
    c.to_host_async(hv, stats.singleton_row_val, m, "singleton tape vals");
    c.to_host_async(hal, plan.new_AL, m, "singleton tape AL");
    c.to_host_async(hau, plan.new_AU, m, "singleton tape AU");
    c.sync("singleton tape");
    for (int32_t i = 0; i < m; ++i)
      if (hr[i]) append_record(plan.tape, RT::DeletedRow, {i, hc[i]}, {hal[i], hau[i], hv[i]}, DM::Minimal);
  }
  check(gpk_singleton_rows_commit(m, n, P, cl, cu, remove, s), "singleton rows commit");
  plan.has_row_action = true;
  plan.has_change = true;
}

// rule_infeasible_redundant_rows.cu:119-243
void apply_rule_infeasible_redundant_rows(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveStatsGpu& stats,
                                          const PresolveParams& prm) {
  amd::AbandonOnThrow abandon_guard;
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t m = lp.A.rows;
  if (m == 0) return;
  auto& c = ctx();
  hipStream_t s = c.stream();
  DBuf<uint8_t> full(m), dlo(m), dhi(m);
  const gpk_plan P = plan_view(plan);
  int32_t* f = c.flags();
  check(gpk_redundant_rows_classify(csr_view(lp.A), P, stats.row_nnz, prm.bound_tol, full, dlo, dhi, f, s),
        "redundant rows classify");
  const int32_t* h = c.read_flags(2);
  if (h[0]) {
    plan.has_infeasible = true;
    return;
  }
  if (!h[1]) return;
  if (prm.record_postsolve_tape && prm.record_postsolve_tape_cpu) {
    std::vector<uint8_t> hf, hl, hh;
    std::vector<double> hal, hau;
    c.begin_copy_group();
    c.to_host_async(hf, full.get(), m, "redundant tape full");
    c.to_host_async(hl, dlo.get(), m, "redundant tape lo");
    c.to_host_async(hh, dhi.get(), m, "redundant tape hi");
    c.to_host_async(hal, plan.new_AL, m, "redundant tape AL");
    c.to_host_async(hau, plan.new_AU, m, "redundant tape AU");
    c.sync("redundant tape");
    const double inf = HUGE_VAL;
    for (int32_t i = 0; i < m; ++i)
      if (hf[i]) append_record(plan.tape, RT::DeletedRow, {i}, {hal[i], hau[i]}, DM::Minimal);
    for (int32_t i = 0; i < m; ++i) {
      if (hl[i]) append_record(plan.tape, RT::LhsChange, {i}, {hal[i], hau[i], -inf, hau[i]}, DM::Minimal);
      if (hh[i]) append_record(plan.tape, RT::RhsChange, {i}, {hal[i], hau[i], hal[i], inf}, DM::Minimal);
    }
  }
  check(gpk_redundant_rows_apply(m, P, full, dlo, dhi, s), "redundant rows apply");
  plan.has_row_action = true;
  plan.has_change = true;
}

// rule_implied_variable_bounds.cu:886-1164
void apply_rule_implied_variable_bounds(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveStatsGpu& stats,
                                        const PresolveParams& prm) {
  amd::AbandonOnThrow abandon_guard;
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0 || n == 0) return;
  auto& c = ctx();
  hipStream_t s = c.stream();
  const gpk_plan P = plan_view(plan);
  const gpk_csr A = csr_view(lp.A), AT = csr_view(lp.AT);
  DBuf<double> rmin(m), rmax(m);
  DBuf<int32_t> rni(m), rpi(m);
  row_activity(A.ptr, A.idx, A.val, m, n, plan.new_l, plan.new_u, prm.zero_tol, rmin, rmax, rni, rpi, s);
  DBuf<gpk_implied_row> rows(m);
  check(gpk_implied_pack_rows(m, P, stats.row_nnz, rmin, rmax, rni, rpi, rows, s), "implied bounds row records");
  DBuf<double> cl(n), cu(n), vmax(n);
  check(gpk_implied_candidates(AT, P, rows, prm.zero_tol, cl, cu, s), "implied bounds candidates");
  check(gpk_col_max_abs(AT.ptr, AT.val, n, vmax, s), "implied bounds col max");
  DBuf<double> fl(n), fu(n), fval(n);
  DBuf<uint8_t> lch(n), uch(n), fixed(n);
  int32_t* f = c.flags();
  check(gpk_implied_finalize(n, P, lp.l ? lp.l : plan.new_l, lp.u ? lp.u : plan.new_u, cl, cu, vmax,
                             prm.implied_variable_bounds_preserve_free_zero_cost ? 1 : 0, prm.zero_tol,
                             prm.feasibility_tol, fl, fu, lch, uch, fixed, fval, f, s),
        "implied bounds finalize");
  const int32_t* h = c.read_flags(3);
  if (h[0]) {
    plan.has_infeasible = true;
    return;
  }
  const bool tightened = h[1] != 0, any_fixed = h[2] != 0;
  if (!(tightened || any_fixed)) return;
  if (prm.record_postsolve_tape && tightened) {
    DBuf<int32_t> sl(n), su(n), cnt(static_cast<int64_t>(n) + 1);
    const double support_tol = std::fmax(10.0 * prm.bound_tol, 1.0e-10);
    check(gpk_implied_support(AT, P, rows, lch, uch, fl, fu, prm.zero_tol, prm.bound_tol, support_tol, sl, su, cnt,
                              s),
          "implied bounds support rows");
    const int32_t records = amd::scan_counts(cnt, n, cnt);
    if (records > 0) {
      DBuf<int32_t> rcol(records), rrow(records);
      DBuf<double> rval(4 * static_cast<int64_t>(records));
      check(gpk_implied_pack_tape(n, P, lch, uch, sl, su, fl, fu, cnt, rcol, rrow, rval, s), "implied bounds pack");
      std::vector<int32_t> hc, hr;
      std::vector<double> hv;
      c.begin_copy_group();
      c.to_host_async(hc, rcol.get(), records, "implied tape cols");
      c.to_host_async(hr, rrow.get(), records, "implied tape rows");
      c.to_host_async(hv, rval.get(), 4 * static_cast<int64_t>(records), "implied tape vals");
      c.sync("implied tape");
      for (int32_t r = 0; r < records; ++r) {
        const double* v = hv.data() + 4 * static_cast<size_t>(r);
        if (hr[r] != INT_MAX)
          append_record(plan.tape, RT::BoundChangeTheRow, {hc[r], hr[r]}, {v[0], v[1], v[2], v[3]}, DM::Minimal);
        else
          append_record(plan.tape, RT::BoundChangeNoRow, {hc[r]}, {v[0], v[1], v[2], v[3]}, DM::Minimal);
      }
    }
  }
  check(gpk_implied_apply(n, P, fl, fu, fixed, s), "implied bounds apply");
  if (any_fixed) {
    check(gpk_row_shift(A, plan.keep_row_mask, fixed, fval, -1.0, plan.new_AL, plan.new_AU, s),
          "implied bounds row shift");
    f = c.flags();
    check(gpk_remove_empty_rows_after_fix(A, P, prm.feasibility_tol, f, s), "implied bounds empty rows");
    if (c.read_flags(4)[3]) {
      plan.has_infeasible = true;
      return;
    }
  }
  const double obj = any_fixed ? amd::fixed_col_tape(prm.record_postsolve_tape ? &plan.tape : nullptr, fixed, nullptr,
                                                     fval, plan.keep_row_mask, plan.new_c, lp.AT, n)
                               : 0.0;
  plan.obj_constant_delta += obj;
  plan.has_row_action = plan.has_row_action || any_fixed;
  plan.has_col_action = plan.has_col_action || any_fixed;
  plan.has_change = true;
}

void release_implied_variable_bounds_workspace() noexcept {}

// rule_duplicate_rows.cu:424-613
void apply_rule_duplicate_rows(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveStatsGpu& stats,
                               const PresolveParams& prm) {
  amd::AbandonOnThrow abandon_guard;
  (void)stats;
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t m = lp.A.rows;
  if (m <= 1) return;
  auto& c = ctx();
  hipStream_t s = c.stream();
  const gpk_csr A = csr_view(lp.A);
  DBuf<uint64_t> hash(m), hash_alt(m);
  DBuf<int32_t> rows(m), rows_alt(m);
  check(gpk_dup_row_hash(A, plan.keep_row_mask, prm.zero_tol, hash, s), "duplicate rows hash");
  check(gpk_iota_i32(rows, m, s), "duplicate rows iota");
  {
    DBuf<int32_t> scratch(gpk_radix_scratch_ints(m));
    check(gpk_radix_sort_u64_i32(hash, rows, hash_alt, rows_alt, m, 64, scratch, s), "duplicate rows sort");
  }
  DBuf<uint8_t> del(m);
  DBuf<int32_t> merge_to(m), step(m);
  DBuf<double> rec(5 * static_cast<int64_t>(m));
  zero(del.get(), m, "duplicate rows delete mask");
  int32_t* f = c.flags();
  check(gpk_dup_row_groups(A, plan_view(plan), hash, rows, prm.zero_tol, prm.feasibility_tol, del, merge_to, step,
                           rec, f, s),
        "duplicate rows groups");
  if (c.read_flags(1)[0]) {
    plan.has_infeasible = true;
    return;
  }
  f = c.flags();
  check(gpk_delete_rows(m, plan.keep_row_mask, del, f, s), "duplicate rows delete");
  if (c.read_flags(1)[0] == 0) return;
  if (prm.record_postsolve_tape) {
    // Only the deleted rows' data is packed on the device and copied back.
    DBuf<int32_t> prefix(static_cast<int64_t>(m) + 1);
    const int32_t k = amd::scan_mask_count(del, m, prefix);
    DBuf<int32_t> d_row(k), d_to(k), d_step(k);
    DBuf<double> d_rec(5 * static_cast<int64_t>(k));
    check(gpk_dup_row_pack(m, del, prefix, merge_to, step, rec, d_row, d_to, d_step, d_rec, s),
          "duplicate rows pack");
    std::vector<int32_t> hrow(k), hto(k), hstep(k);
    std::vector<double> hrec(5 * static_cast<size_t>(k));
    if (k > 0) {
      c.begin_copy_group();
      c.download(hrow.data(), d_row, sizeof(int32_t) * k, "duplicate rows");
      c.download(hto.data(), d_to, sizeof(int32_t) * k, "duplicate rows");
      c.download(hstep.data(), d_step, sizeof(int32_t) * k, "duplicate rows");
      c.download(hrec.data(), d_rec, sizeof(double) * 5 * k, "duplicate rows");
      c.sync("duplicate rows tape");
    }
    // Tape order: by merge step, then by deleted row.
    std::vector<int32_t> order(k);
    for (int32_t t = 0; t < k; ++t) order[t] = t;
    std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
      return hstep[a] != hstep[b] ? hstep[a] < hstep[b] : hrow[a] < hrow[b];
    });
    for (int32_t t : order) {
      const double* v = hrec.data() + 5 * static_cast<size_t>(t);
      append_record(plan.tape, RT::DuplicateRow, {hto[t], hrow[t]}, {v[0], v[1], v[2], v[3], v[4]}, DM::Exact);
    }
  }
  plan.has_row_action = true;
  plan.has_change = true;
}

// rule_infeasible_fixed_variables.cu:117-266
// Detection and decision halves of the fixed-variable and empty-column
// rules.  A detection launch only writes its own scratch and flag words, so
// the scheduler can launch both rules' detections and read their flags in
// one wait when the first rule finds nothing
// (apply_rule_fixed_variables_and_empty_cols).
namespace {

struct ColFixScratch {
  DBuf<uint8_t> fixed;
  DBuf<double> v, contrib;
  explicit ColFixScratch(int32_t n) : fixed(n), v(n), contrib(n) {}
};

void fixed_vars_launch(const PresolvePlanGpu& plan, int32_t n, const PresolveParams& prm, ColFixScratch& w,
                       int32_t* f) {
  check(gpk_fixed_vars_candidates(n, plan_view(plan), prm.bound_tol, w.fixed, w.v, w.contrib, f, ctx().stream()),
        "fixed vars candidates");
}

// h[0]: any fixed column.
void fixed_vars_finish(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveParams& prm, ColFixScratch& w,
                       const int32_t* h) {
  if (h[0] == 0) return;
  const int32_t n = lp.A.cols, m = lp.A.rows;
  auto& c = ctx();
  hipStream_t s = c.stream();
  const gpk_plan P = plan_view(plan);
  if (prm.record_postsolve_tape && prm.record_postsolve_tape_cpu) {
    const std::vector<int32_t> cols = masked_indices(w.fixed, n);
    std::vector<double> hv, hc, av;
    std::vector<uint8_t> kr;
    std::vector<int32_t> ap, ai;
    c.begin_copy_group();
    c.to_host_async(hv, w.v.get(), n, "fixed vars tape v");
    c.to_host_async(hc, plan.new_c, n, "fixed vars tape c");
    c.to_host_async(kr, plan.keep_row_mask, m, "fixed vars tape keep_row");
    c.to_host_async(ap, lp.AT.rowPtr, static_cast<int64_t>(n) + 1, "fixed vars tape AT ptr");
    c.to_host_async(ai, lp.AT.colVal, lp.AT.nnz, "fixed vars tape AT idx");
    c.to_host_async(av, lp.AT.nzVal, lp.AT.nnz, "fixed vars tape AT val");
    c.sync("fixed vars tape");
    for (int32_t j : cols) {
      std::vector<int32_t> idx{j};
      std::vector<double> vals{hv[j], hc[j]};
      for (int32_t q = ap[j]; q < ap[j + 1]; ++q) {
        if (!kr[ai[q]]) continue;
        idx.push_back(ai[q]);
        vals.push_back(av[q]);
      }
      append_record(plan.tape, RT::FixedCol, idx.data(), idx.size(), vals.data(), vals.size(), DM::Minimal);
    }
  }
  // Row shift over every row (the reference does not test keep_row here).
  check(gpk_row_shift(csr_view(lp.A), nullptr, w.fixed, w.v, -1.0, plan.new_AL, plan.new_AU, s),
        "fixed vars row shift");
  check(gpk_fix_cols_apply(n, P, w.fixed, w.v, s), "fixed vars apply");
  amd::add_ordered_masked_sum(&plan.obj_constant_delta, w.fixed, w.contrib, n);
  plan.has_col_action = true;
  plan.has_change = true;
}

void empty_cols_launch(const PresolvePlanGpu& plan, int32_t n, const PresolveStatsGpu& stats,
                       const PresolveParams& prm, ColFixScratch& w, int32_t* f) {
  check(gpk_empty_cols(n, plan_view(plan), stats.empty_col_mask, prm.bound_tol, prm.zero_tol, w.fixed, w.v,
                       w.contrib, f, ctx().stream()),
        "empty cols");
}

// h[0] infeasible, h[1] unbounded, h[2] any fixed column.
void empty_cols_finish(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveParams& prm, ColFixScratch& w,
                       const int32_t* h) {
  if (h[0]) {
    plan.has_infeasible = true;
    return;
  }
  if (h[1]) {
    plan.has_unbounded = true;
    return;
  }
  if (!h[2]) return;
  const int32_t n = lp.A.cols;
  auto& c = ctx();
  if (prm.record_postsolve_tape && prm.record_postsolve_tape_cpu) {
    const std::vector<int32_t> cols = masked_indices(w.fixed, n);
    std::vector<double> hv, hc;
    c.begin_copy_group();
    c.to_host_async(hv, w.v.get(), n, "empty cols tape v");
    c.to_host_async(hc, plan.new_c, n, "empty cols tape c");
    c.sync("empty cols tape");
    for (int32_t j : cols) append_record(plan.tape, RT::FixedCol, {j}, {hv[j], hc[j]}, DM::Minimal);
  }
  check(gpk_fix_cols_apply(n, plan_view(plan), w.fixed, w.v, c.stream()), "empty cols apply");
  amd::add_ordered_masked_sum(&plan.obj_constant_delta, w.fixed, w.contrib, n);
  plan.has_col_action = true;
  plan.has_change = true;
}

}  // namespace

void apply_rule_infeasible_fixed_variables(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveParams& prm) {
  amd::AbandonOnThrow abandon_guard;
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t n = lp.A.cols;
  if (n == 0) return;
  ColFixScratch w(n);
  auto& c = ctx();
  int32_t* f = c.flags();
  fixed_vars_launch(plan, n, prm, w, f);
  fixed_vars_finish(plan, lp, prm, w, c.read_flags(1));
}

// rule_empty_cols.cu:110-198
void apply_rule_empty_cols(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveStatsGpu& stats,
                           const PresolveParams& prm) {
  amd::AbandonOnThrow abandon_guard;
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t n = lp.A.cols;
  if (n == 0) return;
  ColFixScratch w(n);
  auto& c = ctx();
  int32_t* f = c.flags();
  empty_cols_launch(plan, n, stats, prm, w, f);
  empty_cols_finish(plan, lp, prm, w, c.read_flags(3));
}

bool apply_rule_fixed_variables_and_empty_cols(PresolvePlanGpu& plan, const LPInfoGpu& lp,
                                               const PresolveStatsGpu& stats, const PresolveParams& prm) {
  amd::AbandonOnThrow abandon_guard;
  if (plan.has_infeasible || plan.has_unbounded) return true;
  const int32_t n = lp.A.cols;
  if (n == 0) return true;
  ColFixScratch wf(n), we(n);
  auto& c = ctx();
  int32_t* f = c.flags();
  fixed_vars_launch(plan, n, prm, wf, f);
  empty_cols_launch(plan, n, stats, prm, we, f + 4);
  const int32_t* h = c.read_flags(7);
  const int32_t hf[1] = {h[0]}, he[3] = {h[4], h[5], h[6]};
  if (hf[0] != 0) {
    fixed_vars_finish(plan, lp, prm, wf, hf);
    return false;
  }
  empty_cols_finish(plan, lp, prm, we, he);
  return true;
}

// rule_column_singletons.cu:1631-1928
static void column_singletons_mode(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveStatsGpu& stats,
                                   const PresolveParams& prm, bool eq_mode, bool skip_precheck) {
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0 || n == 0) return;
  auto& c = ctx();
  hipStream_t s = c.stream();
  const gpk_plan P = plan_view(plan);
  const gpk_csr A = csr_view(lp.A), AT = csr_view(lp.AT);
  DBuf<uint8_t> needed(m);
  zero(needed.get(), m, "column singletons needed rows");
  int32_t* f = c.flags();
  check(gpk_cs_precheck(n, m, P, stats.column_singleton_mask, stats.column_singleton_row, prm.bound_tol, needed, f, s),
        "column singletons precheck");
  if (!skip_precheck) {
    // Singletons at all; on equality rows (what eq mode uses); on other rows
    // (what dual inference uses).
    const int32_t* h = c.read_flags(3);
    if (!h[0] || (eq_mode && !h[1]) || (!eq_mode && !h[2])) return;
  }
  DBuf<double> fmn(m), fmx(m);
  DBuf<int32_t> mn(m), mp(m), xn(m), xp(m);
  DBuf<uint8_t> invalid(m);
  const gpk_row_cache rc{fmn, fmx, mn, mp, xn, xp, invalid};
  check(gpk_cs_activity_cache(A, P, needed, rc, s), "column singletons activity");
  DBuf<int32_t> owner(m);
  check(gpk_fill_i32(owner, INT_MAX, m, s), "column singletons owner fill");
  // The owner pass's unbounded flag (word 0) and the process pass's three
  // flags (words 4-6) come back in one wait: the process pass writes only
  // this call's buffers, which an unbounded owner pass discards.
  f = c.flags();
  check(gpk_cs_owner(A, P, stats.column_singleton_mask, stats.column_singleton_row, stats.column_singleton_val, rc,
                     prm.bound_tol, prm.zero_tol, eq_mode ? 1 : 0, owner, f, s),
        "column singletons owner");
  DBuf<double> AL_new(m), AU_new(m), side(n), contrib(n), c_new(n);
  DBuf<uint8_t> row_del(m), row_changed(m), col_del(n);
  amd::CopyBatch()
      .zero(row_del.get(), m)
      .zero(row_changed.get(), m)
      .zero(col_del.get(), n)
      .zero(side.get(), n)
      .zero(contrib.get(), n)
      .copy(AL_new.get(), plan.new_AL, m)
      .copy(AU_new.get(), plan.new_AU, m)
      .run("column singletons workspace");
  check(gpk_cs_process(n, m, P, stats.column_singleton_mask, stats.column_singleton_row, stats.column_singleton_val,
                       owner, rc, prm.bound_tol, prm.zero_tol, eq_mode ? 1 : 0, AL_new, AU_new, row_del, row_changed,
                       col_del, side, contrib, f + 4, s),
        "column singletons process");
  const int32_t* h = c.read_flags(7);
  if (h[0] || h[4]) {
    plan.has_unbounded = true;
    return;
  }
  const bool st_eq = h[5] != 0, st_dual = h[6] != 0;
  if (!(st_eq || st_dual)) return;
  if (st_eq && eq_mode)
    check(gpk_cs_objective(AT, P, col_del, owner, stats.column_singleton_val, c_new, s), "column singletons objective");
  else
    d2d(c_new.get(), plan.new_c, n, "cs c copy");
  if (st_eq && eq_mode && prm.record_postsolve_tape) {
    DBuf<uint8_t> sel(m);
    DBuf<int32_t> sel_prefix(static_cast<int64_t>(m) + 1), sup(static_cast<int64_t>(m) + 1);
    check(gpk_cs_subcol_counts(A, P, owner, col_del, sel, sup, s), "subcol counts");
    amd::scan_mask_launch(sel, m, sel_prefix);
    amd::scan_counts_launch(sup, m, sup);
    int32_t records = 0, support = 0;  // both totals in one wait
    c.begin_copy_group();
    c.download(&records, sel_prefix.get() + m, sizeof(int32_t), "subcol record count");
    c.download(&support, sup.get() + m, sizeof(int32_t), "subcol support count");
    c.sync("subcol counts");
    if (records > 0) {
      DBuf<int32_t> oi3(3 * static_cast<int64_t>(records)), scol(support);
      DBuf<double> ov6(6 * static_cast<int64_t>(records)), sval(support);
      check(gpk_cs_subcol_pack(A, P, owner, row_del, side, sel, sel_prefix, sup, oi3, ov6, scol, sval, s),
            "subcol pack");
      // Records are in ascending row order; record r owns the support
      // entries [start_r, start_{r+1}) of the packed arrays.
      std::vector<int32_t> hi3, hsc;
      std::vector<double> hv6, hsv;
      c.begin_copy_group();
      c.to_host_async(hi3, oi3.get(), 3 * static_cast<int64_t>(records), "subcol ids");
      c.to_host_async(hv6, ov6.get(), 6 * static_cast<int64_t>(records), "subcol vals");
      c.to_host_async(hsc, scol.get(), support, "subcol support cols");
      c.to_host_async(hsv, sval.get(), support, "subcol support vals");
      c.sync("subcol tape");
      std::vector<int32_t> idx;
      std::vector<double> vals;
      for (int32_t r = 0; r < records; ++r) {
        const double* v6 = hv6.data() + 6 * static_cast<size_t>(r);
        if (std::fabs(v6[0]) <= prm.zero_tol) continue;
        const int32_t b = hi3[3 * r + 2];
        const int32_t e = r + 1 < records ? hi3[3 * (r + 1) + 2] : support;
        idx.assign({hi3[3 * r], hi3[3 * r + 1], e - b});
        vals.assign(v6, v6 + 6);
        for (int32_t q = b; q < e; ++q) {
          idx.push_back(hsc[q]);
          vals.push_back(hsv[q]);
        }
        append_record(plan.tape, RT::SubCol, idx.data(), idx.size(), vals.data(), vals.size(), DM::Minimal);
      }
    }
  }
  if (st_dual && !eq_mode && prm.record_postsolve_tape) {
    const std::vector<int32_t> rows = masked_indices(row_changed, m);
    for (int32_t row : rows) append_record(plan.tape, RT::EqToIneq, {row}, {0.0}, DM::Minimal);
  }
  check(gpk_cs_apply(m, n, P, row_del, col_del, AL_new, AU_new, c_new, s), "column singletons apply");
  amd::add_ordered_masked_sum(&plan.obj_constant_delta, col_del, contrib, n);
  plan.has_row_action = true;
  plan.has_col_action = plan.has_col_action || st_eq;
  plan.has_change = true;
}

void apply_rule_column_singletons_eq(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveStatsGpu& stats,
                                     const PresolveParams& prm, bool skip_precheck) {
  amd::AbandonOnThrow abandon_guard;
  column_singletons_mode(plan, lp, stats, prm, true, skip_precheck);
}

void apply_rule_column_singletons_dual_infer(PresolvePlanGpu& plan, const LPInfoGpu& lp,
                                             const PresolveStatsGpu& stats, const PresolveParams& prm) {
  amd::AbandonOnThrow abandon_guard;
  column_singletons_mode(plan, lp, stats, prm, false, false);
}

void release_column_singletons_workspace() noexcept {}


// ---------------------------------------------------------------------------
// Doubleton equations (rule_doubleton_equations.cu)
namespace {

constexpr int kDblMatchRounds = 8;
constexpr int kDblInnerRounds = 1024;

struct DblWork {
  DBuf<uint8_t> acceptable, active, selected, blocked;
  DBuf<int32_t> cec, ckc, live_nnz;
  DBuf<uint64_t> owner;
  DBuf<double> cev, ckv;
  DBuf<uint8_t> is_elim, is_fixed_keep;
  DBuf<double> alpha_of, beta_of, fixed_val_of;
  DBuf<int32_t> row_of, keep_of;
  DblWork(int32_t m, int32_t n)
      : acceptable(m), active(m), selected(m), blocked(n), cec(m), ckc(m), live_nnz(n), owner(n), cev(m), ckv(m),
        is_elim(n), is_fixed_keep(n), alpha_of(n), beta_of(n), fixed_val_of(n), row_of(n), keep_of(n) {}
  gpk_dbl view() {
    return gpk_dbl{acceptable, active, selected, blocked, cec, ckc, cev, ckv, owner, live_nnz};
  }
  gpk_dbl_cols cols() { return gpk_dbl_cols{is_elim, is_fixed_keep, alpha_of, beta_of, fixed_val_of, row_of, keep_of}; }
};

struct DblTapeRec {
  int32_t i3[3];
  double v12[12];
};

// One batch (rule_doubleton_equations.cu:1432-1677).  Returns true when the
// plan changed; os/of receive the substitution and fixed-pair objective
// deltas of the batch.
// The transpose of A + delta, built from AT: the delta (rows sorted by
// column) is transposed (it is small), then every row of AT is merged with
// its row of the transposed delta by k_dbl_merge, the kernel that builds the
// new A.  Per entry the same terms are added in the same order (the old
// coefficient, then the delta terms in row-major delta order, which a stable
// transpose keeps) and the same |value| <= zero_tol test drops it, so the
// result is bit for bit the transpose of the merged A; nnz is known (the new
// A's), so no count is read back.
DeviceCsrMatrix merged_transpose(const DeviceCsrMatrix& AT, const int32_t* dptr, const int32_t* dcol,
                                 const double* dval, int32_t total, int32_t m, double zt, int32_t nnz) {
  auto& c = ctx();
  hipStream_t s = c.stream();
  const int32_t n = AT.rows;
  const DeviceCsrMatrix delta{m, n, total, const_cast<int32_t*>(dptr), const_cast<int32_t*>(dcol),
                              const_cast<double*>(dval)};
  DeviceCsrMatrix dT = amd::transpose_csr(delta);
  DeviceCsrMatrix out;
  out.rows = n;
  out.cols = m;
  out.nnz = nnz;
  try {
    DBuf<int32_t> tptr(static_cast<int64_t>(n) + 1);
    check(gpk_dbl_merge(csr_view(AT), dT.rowPtr, dT.colVal, dT.nzVal, zt, nullptr, tptr, nullptr, nullptr, s),
          "doubleton AT merge count");
    amd::scan_counts_launch(tptr, n, tptr);
    DBuf<int32_t> tidx(nnz);
    DBuf<double> tval(nnz);
    check(gpk_dbl_merge(csr_view(AT), dT.rowPtr, dT.colVal, dT.nzVal, zt, tptr, nullptr, tidx, tval, s),
          "doubleton AT merge");
    out.rowPtr = tptr.release();
    out.colVal = tidx.release();
    out.nzVal = tval.release();
  } catch (...) {
    amd::free_csr(dT);
    throw;
  }
  amd::free_csr(dT);
  return out;
}

// at_out (when not null and A is rebuilt): receives the transpose of the new
// A, formed as AT_src merged with the transposed delta.  AT_src must then be
// the exact transpose of this round's A with sorted rows.
bool dbl_batch_once(PresolvePlanGpu& plan, const LPInfoGpu& lp, const DeviceCsrMatrix& AT_src,
                    const PresolveParams& prm, DblWork& w, std::vector<DblTapeRec>* recs, double* os, double* of,
                    DeviceCsrMatrix* at_out) {
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0 || n == 0) return false;
  auto& c = ctx();
  hipStream_t s = c.stream();
  const gpk_plan P = plan_view(plan);
  const gpk_csr A = csr_view(plan.has_new_A ? plan.new_A : lp.A);
  const gpk_csr AT = csr_view(AT_src);
  const double zt = prm.zero_tol, tol = prm.bound_tol;
  const double rtol = std::fmax(1.0e-9, prm.bound_tol);
  gpk_dbl d = w.view();
  check(gpk_dbl_live_nnz(AT, P, zt, d.live_nnz, s), "doubleton live nnz");
  int32_t* f = c.flags();
  check(gpk_dbl_candidates(A, AT, P, d, zt, tol, rtol, prm.doubleton_equations_max_fill_in_proxy, 1, f, s),
        "doubleton candidates");
  amd::CopyBatch().zero(d.blocked, n).zero(d.owner, n).run("doubleton blocked and owners");
  for (int r = 1; r <= kDblMatchRounds; ++r) {
    check(gpk_dbl_claim(m, d, r, s), "doubleton claim");
    check(gpk_dbl_select(m, d, r, f, s), "doubleton select");
  }
  const int32_t* h = c.read_flags(2);
  const int32_t acceptable = h[0], k = h[1];
  const double ratio = acceptable == 0 ? 0.0 : static_cast<double>(k) / acceptable;
  if (k == 0 || (prm.doubleton_equations_scan && (k < prm.doubleton_equations_min_selected_per_batch ||
                                                  ratio < prm.doubleton_equations_min_selected_ratio)))
    return false;
  DBuf<int32_t> prefix(static_cast<int64_t>(m) + 1), sel(k);
  amd::scan_mask_count(d.selected, m, prefix);
  check(gpk_pack_index(d.selected, prefix, m, sel, s), "doubleton pack selected");
  gpk_dbl_cols cols = w.cols();
  amd::CopyBatch().zero(cols.is_elim, n).zero(cols.is_fixed_keep, n).run("doubleton column flags");
  DBuf<uint8_t> subst(k), fixedp(k);
  DBuf<double> osub(k), ofix(k), rv(12 * static_cast<int64_t>(k));
  DBuf<int32_t> ri(3 * static_cast<int64_t>(k));
  f = c.flags();
  check(gpk_dbl_apply(k, sel, P, d, cols, tol, subst, fixedp, osub, ofix, ri, rv, f, s), "doubleton apply");
  h = c.read_flags(4);
  const bool infeasible = h[0] != 0;
  const int32_t subst_count = h[2], fixed_count = h[3];
  *os = amd::ordered_masked_sum(subst, osub, k);
  *of = amd::ordered_masked_sum(fixedp, ofix, k);
  if (infeasible) {
    plan.has_infeasible = true;
    return false;
  }
  if (recs != nullptr && subst_count > 0) {
    std::vector<uint8_t> hs;
    std::vector<int32_t> hi;
    std::vector<double> hv;
    c.begin_copy_group();
    c.to_host_async(hs, subst.get(), k, "doubleton tape mask");
    c.to_host_async(hi, ri.get(), 3 * static_cast<int64_t>(k), "doubleton tape ids");
    c.to_host_async(hv, rv.get(), 12 * static_cast<int64_t>(k), "doubleton tape vals");
    c.sync("doubleton tape");
    for (int32_t t = 0; t < k; ++t) {
      if (!hs[t]) continue;
      DblTapeRec r;
      std::copy(hi.begin() + 3 * t, hi.begin() + 3 * t + 3, r.i3);
      std::copy(hv.begin() + 12 * static_cast<size_t>(t), hv.begin() + 12 * static_cast<size_t>(t) + 12, r.v12);
      recs->push_back(r);
    }
  }
  int32_t total = 0;
  DBuf<int32_t> dptr;
  if (subst_count > 0) {
    dptr.alloc(static_cast<int64_t>(m) + 1);
    check(gpk_dbl_delta_count(A, P, cols, zt, dptr, s), "doubleton delta count");
    total = amd::scan_counts(dptr, m, dptr);
  }
  // Row shift: fixed keep columns, then substituted columns when A is rebuilt.
  if (fixed_count > 0 || total > 0)
    check(gpk_dbl_shift(A, P, cols, zt, total > 0 ? 1 : 0, s), "doubleton row shift");
  if (total > 0) {
    DBuf<int32_t> dcol(total);
    DBuf<double> dval(total);
    check(gpk_dbl_delta_fill(A, P, cols, zt, dptr, dcol, dval, s), "doubleton delta fill");
    DeviceCsrMatrix out;
    out.rows = A.rows;
    out.cols = A.cols;
    DBuf<int32_t> optr(static_cast<int64_t>(m) + 1);
    check(gpk_dbl_merge(A, dptr, dcol, dval, zt, nullptr, optr, nullptr, nullptr, s), "doubleton merge count");
    out.nnz = amd::scan_counts(optr, m, optr);
    DBuf<int32_t> oidx(out.nnz);
    DBuf<double> oval(out.nnz);
    check(gpk_dbl_merge(A, dptr, dcol, dval, zt, optr, nullptr, oidx, oval, s), "doubleton merge");
    if (at_out) *at_out = merged_transpose(AT_src, dptr, dcol, dval, total, m, zt, out.nnz);
    out.rowPtr = optr.release();
    out.colVal = oidx.release();
    out.nzVal = oval.release();
    if (plan.has_new_A) amd::free_csr(plan.new_A);
    plan.new_A = out;
    plan.has_new_A = true;
  }
  plan.has_row_action = subst_count > 0;
  plan.has_col_action = true;
  plan.has_change = true;
  return true;
}

// Single-doubleton path, doubleton_equations_scan = false
// (rule_doubleton_equations.cu:1913-2084, 266-445, 1779-1909).
void dbl_scalar(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveParams& prm) {
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0 || n == 0) return;
  auto& c = ctx();
  hipStream_t s = c.stream();
  const gpk_plan P = plan_view(plan);
  const gpk_csr A = csr_view(plan.has_new_A ? plan.new_A : lp.A);
  const gpk_csr AT = csr_view(lp.AT);
  const double zt = prm.zero_tol, tol = prm.bound_tol;
  const double rtol = std::fmax(1.0e-9, prm.bound_tol);
  DblWork w(m, n);
  gpk_dbl d = w.view();
  check(gpk_dbl_live_nnz(AT, P, zt, d.live_nnz, s), "doubleton live nnz");
  int32_t* f = c.flags();
  check(gpk_dbl_candidates(A, AT, P, d, zt, tol, rtol, prm.doubleton_equations_max_fill_in_proxy, 0, f, s),
        "doubleton candidates");
  if (c.read_flags(1)[0] == 0) return;
  const std::vector<int32_t> rows = masked_indices(d.acceptable, m);
  const int32_t target = rows.front();  // smallest row index (atomicMin in the reference)
  const int32_t ec = c.scalar_to_host(d.cec + target, "doubleton elim col");
  const int32_t kc = c.scalar_to_host(d.ckc + target, "doubleton keep col");
  const double ev = c.scalar_to_host(d.cev + target, "doubleton elim val");
  const double kv = c.scalar_to_host(d.ckv + target, "doubleton keep val");
  const double rhs = c.scalar_to_host(plan.new_AU + target, "doubleton rhs");
  const double ol_e = c.scalar_to_host(plan.new_l + ec, "doubleton l_e");
  const double ou_e = c.scalar_to_host(plan.new_u + ec, "doubleton u_e");
  const double oc_e = c.scalar_to_host(plan.new_c + ec, "doubleton c_e");
  const double ol_k = c.scalar_to_host(plan.new_l + kc, "doubleton l_k");
  const double ou_k = c.scalar_to_host(plan.new_u + kc, "doubleton u_k");
  const double oc_k = c.scalar_to_host(plan.new_c + kc, "doubleton c_k");
  const double al = -kv / ev, be = rhs / ev;
  double ml = -HUGE_VAL, mu = HUGE_VAL;
  if (std::isfinite(ol_e)) {
    const double b = (ol_e - be) / al;
    if (al > 0.0) ml = b; else mu = b;
  }
  if (std::isfinite(ou_e)) {
    const double b = (ou_e - be) / al;
    if (al > 0.0) mu = b; else ml = b;
  }
  if (ml > mu) std::swap(ml, mu);
  const double nkl = std::isfinite(ml) ? std::fmax(ol_k, ml) : ol_k;
  const double nku = std::isfinite(mu) ? std::fmin(ou_k, mu) : ou_k;
  if (nkl > nku + tol) {
    plan.has_infeasible = true;
    return;
  }
  const int32_t ec_b = c.scalar_to_host(lp.AT.rowPtr + ec, "doubleton AT ptr");
  const int32_t ec_e = c.scalar_to_host(lp.AT.rowPtr + ec + 1, "doubleton AT ptr");
  check(gpk_dbl_scalar_shift(AT, ec, ec_e - ec_b, be, plan.new_AL, plan.new_AU, s), "doubleton scalar shift");
  check(gpk_fill_f64(plan.new_l + kc, nkl, 1, s), "doubleton set l");
  check(gpk_fill_f64(plan.new_u + kc, nku, 1, s), "doubleton set u");
  check(gpk_fill_f64(plan.new_c + kc, oc_k + oc_e * al, 1, s), "doubleton set c");
  check(gpk_fill_u8(plan.keep_row_mask + target, 0, 1, s), "doubleton drop row");
  check(gpk_fill_u8(plan.keep_col_mask + ec, 0, 1, s), "doubleton drop col");
  if (prm.record_postsolve_tape)
    append_record(plan.tape, RT::DoubletonEquation, {ec, kc, target, 0},
                  {ev, kv, rhs, ol_e, ou_e, ol_k, ou_k, nkl, nku, oc_e, 0.0, 0.0}, DM::Minimal);
  DeviceCsrMatrix out;
  out.rows = A.rows;
  out.cols = A.cols;
  DBuf<int32_t> optr(static_cast<int64_t>(m) + 1);
  check(gpk_dbl_scalar_rows(A, target, ec, kc, al, zt, nullptr, optr, nullptr, nullptr, s), "doubleton rows count");
  out.nnz = amd::scan_counts(optr, m, optr);
  DBuf<int32_t> oidx(out.nnz);
  DBuf<double> oval(out.nnz);
  check(gpk_dbl_scalar_rows(A, target, ec, kc, al, zt, optr, nullptr, oidx, oval, s), "doubleton rows");
  out.rowPtr = optr.release();
  out.colVal = oidx.release();
  out.nzVal = oval.release();
  if (plan.has_new_A) amd::free_csr(plan.new_A);
  plan.new_A = out;
  plan.has_new_A = true;
  plan.obj_constant_delta += oc_e * be;
  plan.has_row_action = true;
  plan.has_col_action = true;
  plan.has_change = true;
}

}  // namespace

// rule_doubleton_equations.cu:1679-1777 (batch driver) and 1913-2084
void apply_rule_doubleton_equations(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveStatsGpu& stats,
                                    const PresolveParams& prm) {
  (void)stats;
  DeviceCsrMatrix at;
  apply_rule_doubleton_equations_with_at(plan, lp, prm, &at);
  amd::free_csr(at);
}

void apply_rule_doubleton_equations_with_at(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveParams& prm,
                                            DeviceCsrMatrix* new_at) {
  amd::AbandonOnThrow abandon_guard;
  *new_at = DeviceCsrMatrix{};
  if (plan.has_infeasible || plan.has_unbounded) return;
  if (!prm.doubleton_equations_scan) {
    dbl_scalar(plan, lp, prm);
    return;
  }
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0 || n == 0) return;
  amd::sort_rows_if_needed(lp.AT);
  // The AT of each round's A: the working AT while A is the working A, then
  // the merged transpose of every rebuilt A (a plan that arrives with a
  // rebuilt A has no AT of it: those rounds transpose).
  const bool merge = !plan.has_new_A;
  DeviceCsrMatrix at_own;
  bool owns_at = false;
  const DeviceCsrMatrix* at = &lp.AT;
  const int limit = prm.doubleton_equations_max_batch_rounds > 0 &&
                            prm.doubleton_equations_max_batch_rounds < kDblInnerRounds
                        ? prm.doubleton_equations_max_batch_rounds
                        : kDblInnerRounds;
  const bool time_limit = prm.doubleton_equations_max_time > 0.0 && std::isfinite(prm.doubleton_equations_max_time);
  const auto start = std::chrono::steady_clock::now();
  std::vector<DblTapeRec> recs;
  std::vector<double> objs;
  DblWork w(m, n);
  try {
    for (int round = 0; round < limit; ++round) {
      if (time_limit) {
        const std::chrono::duration<double> el = std::chrono::steady_clock::now() - start;
        if (el.count() >= prm.doubleton_equations_max_time) break;
      }
      double os = 0.0, of = 0.0;
      DeviceCsrMatrix merged;
      const int32_t* a_before = plan.has_new_A ? plan.new_A.rowPtr : nullptr;
      const bool changed = dbl_batch_once(plan, lp, *at, prm, w, prm.record_postsolve_tape ? &recs : nullptr, &os, &of,
                                          merge ? &merged : nullptr);
      if (plan.has_infeasible || plan.has_unbounded || !changed) {
        amd::free_csr(merged);
        break;
      }
      objs.push_back(os);
      objs.push_back(of);
      if (plan.has_new_A && plan.new_A.rowPtr != a_before) {  // A rebuilt this round
        DeviceCsrMatrix t = merged.rowPtr ? merged : amd::transpose_csr(plan.new_A);
        if (owns_at) amd::free_csr(at_own);
        at_own = t;
        owns_at = true;
        at = &at_own;
      }
    }
  } catch (...) {
    if (owns_at) amd::free_csr(at_own);
    throw;
  }
  if (owns_at) {
    if (plan.has_new_A && !plan.has_infeasible && !plan.has_unbounded) *new_at = at_own;
    else amd::free_csr(at_own);
  }
  for (const DblTapeRec& r : recs)
    append_record(plan.tape, RT::DoubletonEquation, {r.i3[0], r.i3[1], r.i3[2], 0}, {r.v12[0], r.v12[1], r.v12[2],
                  r.v12[3], r.v12[4], r.v12[5], r.v12[6], r.v12[7], r.v12[8], r.v12[9], r.v12[10], r.v12[11]},
                  DM::Minimal);
  for (double o : objs) plan.obj_constant_delta += o;
}

// ---------------------------------------------------------------------------
// rule_dual_fix.cu:174-554
void apply_rule_dual_fix(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveStatsGpu& stats,
                         const PresolveParams& prm, const std::uint8_t* candidate_col_mask) {
  amd::AbandonOnThrow abandon_guard;
  (void)stats;
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t n = lp.A.cols, m = lp.A.rows;
  if (n == 0) return;
  auto& c = ctx();
  hipStream_t s = c.stream();
  const gpk_plan P = plan_view(plan);
  DBuf<uint8_t> fixed(n), inf_fix(n), finite_fix(n);
  DBuf<double> v(n), contrib(n);
  int32_t* f = c.flags();
  check(gpk_dual_fix_candidates(csr_view(lp.AT), P, candidate_col_mask, prm.zero_tol, prm.bound_tol,
                                prm.dual_fix_max_nonzero_cost_col_degree, fixed, inf_fix, finite_fix, v, contrib, f, s),
        "dual fix candidates");
  const int32_t* h = c.read_flags(3);
  if (h[0] == 1) {
    plan.has_infeasible = true;
    return;
  }
  if (h[0] == 2) {
    plan.has_unbounded = true;
    return;
  }
  if (!h[1]) return;
  const bool any_inf = h[2] != 0;
  DBuf<double> shift(m);
  DBuf<uint8_t> row_remove(m);
  f = c.flags();
  if (m > 0)
    check(gpk_dual_fix_rows(csr_view(lp.A), P, finite_fix, inf_fix, v, shift, row_remove, f, s), "dual fix rows");
  const bool row_changed = m > 0 && c.read_flags(4)[3] != 0;
  if (prm.record_postsolve_tape) {
    if (any_inf) {
      // FixedColInf: the column, its infinite side and every live row it
      // touches with that row's entries (rule_dual_fix.cu:386-470).
      const std::vector<int32_t> cols = masked_indices(inf_fix, n);
      std::vector<double> hv, hl, hu, hal, hau, av;
      std::vector<uint8_t> kr;
      std::vector<int32_t> tp, ti, ap, ai;
      c.begin_copy_group();
      c.to_host_async(hv, v.get(), n, "dual fix tape v");
      c.to_host_async(hl, plan.new_l, n, "dual fix tape l");
      c.to_host_async(hu, plan.new_u, n, "dual fix tape u");
      c.to_host_async(kr, plan.keep_row_mask, m, "dual fix tape keep_row");
      c.to_host_async(hal, plan.new_AL, m, "dual fix tape AL");
      c.to_host_async(hau, plan.new_AU, m, "dual fix tape AU");
      c.to_host_async(tp, lp.AT.rowPtr, static_cast<int64_t>(n) + 1, "dual fix tape AT ptr");
      c.to_host_async(ti, lp.AT.colVal, lp.AT.nnz, "dual fix tape AT idx");
      c.to_host_async(ap, lp.A.rowPtr, static_cast<int64_t>(m) + 1, "dual fix tape A ptr");
      c.to_host_async(ai, lp.A.colVal, lp.A.nnz, "dual fix tape A idx");
      c.to_host_async(av, lp.A.nzVal, lp.A.nnz, "dual fix tape A val");
      c.sync("dual fix tape");
      std::vector<int32_t> idx;
      std::vector<double> vals;
      for (int32_t j : cols) {
        const bool pos = hv[j] > 0.0;
        idx.assign({pos ? 1 : -1, j});
        vals.assign({0.0, pos ? hl[j] : hu[j]});
        int32_t active = 0;
        for (int32_t q = tp[j]; q < tp[j + 1]; ++q) {
          const int32_t row = ti[q];
          if (!kr[row]) continue;
          ++active;
          idx.push_back(ap[row + 1] - ap[row]);
          vals.push_back(std::isfinite(hal[row]) ? hal[row] : hau[row]);
          for (int32_t t = ap[row]; t < ap[row + 1]; ++t) {
            idx.push_back(ai[t]);
            vals.push_back(av[t]);
          }
        }
        vals[0] = static_cast<double>(active);
        append_record(plan.tape, RT::FixedColInf, idx.data(), idx.size(), vals.data(), vals.size(), DM::Exact);
      }
    }
    if (prm.record_postsolve_tape_cpu) {
      DBuf<uint8_t> keep_after(m);
      d2d(keep_after.get(), plan.keep_row_mask, m, "dual fix keep_after");
      f = c.flags();
      if (m > 0) check(gpk_delete_rows(m, keep_after, row_remove, f, s), "dual fix keep_after");
      amd::fixed_col_tape(&plan.tape, fixed, inf_fix, v, keep_after, plan.new_c, lp.AT, n);
    }
  }
  check(gpk_dual_fix_apply(m, n, P, fixed, v, shift, row_remove, s), "dual fix apply");
  amd::add_ordered_masked_sum(&plan.obj_constant_delta, fixed, contrib, n);
  plan.has_row_action = plan.has_row_action || row_changed;
  plan.has_col_action = true;
  plan.has_change = true;
}

// ---------------------------------------------------------------------------
// rule_duplicate_columns.cu:95-842
void apply_rule_duplicate_columns(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveStatsGpu& stats,
                                  const PresolveParams& prm) {
  amd::AbandonOnThrow abandon_guard;
  (void)stats;
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t n = lp.A.cols, m = lp.A.rows;
  if (n <= 1) return;
  auto& c = ctx();
  hipStream_t s = c.stream();
  if (lp.AT.nnz > 0) amd::sort_rows_if_needed(lp.AT);
  const gpk_plan P = plan_view(plan);
  const gpk_csr AT = csr_view(lp.AT);
  const double zt = prm.zero_tol;
  const double ctol = std::fmax(prm.zero_tol, prm.bound_tol);
  const double otol = ctol;
  DBuf<uint64_t> hash(n), hash_alt(n);
  DBuf<int32_t> cols(n), cols_alt(n);
  check(gpk_dup_col_hash(AT, P, zt, hash, s), "duplicate cols hash");
  check(gpk_iota_i32(cols, n, s), "duplicate cols iota");
  {
    DBuf<int32_t> scratch(gpk_radix_scratch_ints(n));
    check(gpk_radix_sort_u64_i32(hash, cols, hash_alt, cols_alt, n, 64, scratch, s), "duplicate cols sort");
  }
  const double max_ratio = prm.duplicate_columns_max_abs_ratio;
  const bool use_guard = std::isfinite(max_ratio) && prm.duplicate_columns_scale_guard_min_groups > 1;
  DBuf<int32_t> guard(1);
  if (use_guard) {
    zero(guard.get(), 1, "duplicate cols guard count");
    check(gpk_dup_col_guard(AT, P, hash, cols, zt, ctol, otol, max_ratio,
                            prm.duplicate_columns_scale_guard_max_group_size, guard, s),
          "duplicate cols guard");
  }
  DBuf<uint8_t> del(n), fixed(n);
  DBuf<double> fixed_val(n), rec5(5 * static_cast<int64_t>(n)), contrib(n);
  DBuf<int32_t> merge_to(n);
  amd::CopyBatch().zero(del.get(), n).zero(fixed.get(), n).run("duplicate cols masks");
  int32_t* f = c.flags();
  check(gpk_dup_col_groups(AT, P, hash, cols, zt, ctol, otol, max_ratio, use_guard ? guard.get() : nullptr,
                           prm.duplicate_columns_scale_guard_min_groups, del, fixed, fixed_val, merge_to, rec5, f, s),
        "duplicate cols groups");
  const int32_t* h = c.read_flags(3);
  if (h[0]) {
    plan.has_unbounded = true;
    return;
  }
  const bool merged = h[1] != 0, any_fixed = h[2] != 0;
  if (!(merged || any_fixed)) return;
  if (merged && prm.record_postsolve_tape) {
    const std::vector<int32_t> deleted = masked_indices(del, n);
    std::vector<int32_t> hto;
    std::vector<double> hr;
    c.begin_copy_group();
    c.to_host_async(hto, merge_to.get(), n, "duplicate cols merge_to");
    c.to_host_async(hr, rec5.get(), 5 * static_cast<int64_t>(n), "duplicate cols values");
    c.sync("duplicate cols tape");
    for (int32_t k : deleted) {
      const double* r = hr.data() + 5 * static_cast<size_t>(k);
      append_record(plan.tape, RT::DuplicateColumn, {k, hto[k]}, {r[0], r[1], r[2], r[3], r[4]}, DM::Minimal);
    }
  }
  if (any_fixed) {
    if (prm.record_postsolve_tape)
      amd::fixed_col_tape(&plan.tape, fixed, nullptr, fixed_val, plan.keep_row_mask, plan.new_c, lp.AT, n);
    if (m > 0)
      check(gpk_row_shift(csr_view(lp.A), plan.keep_row_mask, fixed, fixed_val, -1.0, plan.new_AL, plan.new_AU, s),
            "duplicate cols row shift");
    plan.has_row_action = true;
  }
  check(gpk_dup_col_apply(n, P, del, fixed, fixed_val, contrib, s), "duplicate cols apply");
  amd::add_ordered_masked_sum(&plan.obj_constant_delta, fixed, contrib, n);
  plan.has_col_action = true;
  plan.has_change = true;
}

// ---------------------------------------------------------------------------
// rule_redundant_bounds.cu:31-411 (opt-in)
void apply_rule_redundant_bounds(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveStatsGpu& stats,
                                 const PresolveParams& prm) {
  amd::AbandonOnThrow abandon_guard;
  (void)stats;
  if (plan.has_infeasible || plan.has_unbounded) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (m == 0 || n == 0) return;
  auto& c = ctx();
  hipStream_t s = c.stream();
  const gpk_plan P = plan_view(plan);
  const gpk_csr A = csr_view(lp.A), AT = csr_view(lp.AT);
  DBuf<uint8_t> dlo(n), dhi(n), sl(n), su(n);
  DBuf<uint64_t> rkey(m);
  bool changed = false;
  for (;;) {
    check(gpk_rb_candidates(A, AT, P, lp.c, prm.enable_redundant_bounds ? 1 : 0, prm.zero_tol, prm.bound_tol, dlo,
                            dhi, s),
          "redundant bounds candidates");
    int32_t* f = c.flags();
    check(gpk_rb_select(A, AT, dlo, dhi, rkey, sl, su, f, s), "redundant bounds select");
    if (c.read_flags(1)[0] == 0) break;
    if (prm.record_postsolve_tape) {
      std::vector<uint8_t> hsl, hsu;
      std::vector<double> hl, hu;
      c.begin_copy_group();
      c.to_host_async(hsl, sl.get(), n, "redundant bounds lo");
      c.to_host_async(hsu, su.get(), n, "redundant bounds hi");
      c.to_host_async(hl, plan.new_l, n, "redundant bounds l");
      c.to_host_async(hu, plan.new_u, n, "redundant bounds u");
      c.sync("redundant bounds tape");
      for (int32_t j = 0; j < n; ++j) {
        if (!hsl[j] && !hsu[j]) continue;
        const double nl = hsl[j] ? -HUGE_VAL : hl[j];
        const double nu = hsu[j] ? HUGE_VAL : hu[j];
        append_record(plan.tape, RT::BoundChangeNoRow, {j}, {hl[j], hu[j], nl, nu}, DM::Minimal);
      }
    }
    check(gpk_rb_apply(n, P, sl, su, s), "redundant bounds apply");
    changed = true;
  }
  if (changed) {
    plan.has_col_action = true;
    plan.has_change = true;
  }
}

}  // namespace gpu_presolver::presolve
