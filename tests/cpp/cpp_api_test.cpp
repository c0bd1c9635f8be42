// A C++ caller of the library's source-compatible API
// (include/gpu_presolver/presolve/*.hpp), compiled against the headers and
// linked to libgpu_presolver_amd.so the way an existing user of the
// reference API would be.  Each case prints "ok <name>"; any failed check
// prints "FAIL <name>: <what>" and the program exits with status 1.
//
// Cases (the reference's C++-level tests they follow):
//  - summary ownership: move-only, move-assign releases the target's model,
//    self-move keeps it, borrowed models are never freed, free is idempotent
//    (cpp/tests/test_gpu_postsolve.cpp:1152-1189);
//  - presolve of LPs allocated with hipMalloc through
//    run_gpu_presolve_with_record / _with_reduced_lp, postsolve_gpu with the
//    original model, outputs released with hipFree
//    (cpp/tests/test_gpu_direct_port.cpp:22-120 and :273-330);
//  - rule calls on plans that set only the vectors the rule reads, the rest
//    left null, and an LP whose CSR arrays are null
//    (cpp/tests/test_rule_empty_rows.cpp:20-56, test_rule_empty_cols.cpp).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "gpu_presolver/presolve/gpu_postsolve.hpp"
#include "gpu_presolver/presolve/gpu_presolve.hpp"
#include "gpu_presolver/presolve/rules/rule_empty_cols.hpp"
#include "gpu_presolver/presolve/rules/rule_empty_rows.hpp"
#include "gpu_presolver/presolve/rules/rule_singleton_rows.hpp"

namespace gp = gpu_presolver::presolve;

namespace {

int g_failures = 0;
const char* g_case = "";

void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAIL %s: %s\n", g_case, what);
    ++g_failures;
  }
}

void hip_ok(hipError_t st, const char* what) {
  if (st != hipSuccess) {
    std::printf("FAIL %s: %s: %s\n", g_case, what, hipGetErrorString(st));
    std::exit(1);
  }
}

template <class T>
T* dev(const std::vector<T>& v) {
  T* p = nullptr;
  hip_ok(hipMalloc(&p, sizeof(T) * (v.empty() ? 1 : v.size())), "hipMalloc");
  if (!v.empty()) hip_ok(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), "hipMemcpy H2D");
  return p;
}

template <class T>
std::vector<T> host(const T* p, size_t n) {
  std::vector<T> v(n);
  if (n) hip_ok(hipMemcpy(v.data(), p, sizeof(T) * n, hipMemcpyDeviceToHost), "hipMemcpy D2H");
  return v;
}

void free_lp(gp::LPInfoGpu& lp) {
  for (void* p : {static_cast<void*>(lp.A.rowPtr), static_cast<void*>(lp.A.colVal), static_cast<void*>(lp.A.nzVal),
                  static_cast<void*>(lp.AT.rowPtr), static_cast<void*>(lp.AT.colVal), static_cast<void*>(lp.AT.nzVal),
                  static_cast<void*>(lp.c), static_cast<void*>(lp.l), static_cast<void*>(lp.u),
                  static_cast<void*>(lp.AL), static_cast<void*>(lp.AU)})
    if (p) (void)hipFree(p);
  lp = gp::LPInfoGpu{};
}

const double kInf = std::numeric_limits<double>::infinity();

// ---------------------------------------------------------------------------
void summary_ownership() {
  g_case = "summary_ownership";
  static_assert(!std::is_copy_constructible_v<gp::GpuPresolveSummary>, "summary must not be copyable");
  static_assert(!std::is_copy_assignable_v<gp::GpuPresolveSummary>, "summary must not be copy-assignable");
  static_assert(std::is_nothrow_move_constructible_v<gp::GpuPresolveSummary>, "noexcept move");
  static_assert(std::is_nothrow_move_assignable_v<gp::GpuPresolveSummary>, "noexcept move-assign");
  auto owned = [](double cval) {
    gp::GpuPresolveSummary s;
    s.original_cols = s.reduced_cols = 1;
    s.reduced_lp.c = dev<double>({cval});
    s.reduced_lp.A.rowPtr = dev<int32_t>({0});
    s.record.col_red2org = {0};
    s.owns_reduced_lp = true;
    return s;
  };
  gp::GpuPresolveSummary a = owned(3.0);
  gp::GpuPresolveSummary b = std::move(a);
  expect(!a.owns_reduced_lp && a.reduced_lp.c == nullptr, "moved-from summary gave up its model");
  expect(b.owns_reduced_lp && b.reduced_lp.c != nullptr, "moved-to summary owns the model");
  gp::free_gpu_presolve_reduced_lp(a);  // nothing to free
  gp::GpuPresolveSummary t = owned(5.0);
  t = std::move(b);  // releases t's own model first
  expect(!b.owns_reduced_lp && b.reduced_lp.A.rowPtr == nullptr, "move-assign source emptied");
  expect(host(t.reduced_lp.c, 1)[0] == 3.0, "move-assign target holds the moved model");
  gp::GpuPresolveSummary& alias = t;
  t = std::move(alias);
  expect(t.owns_reduced_lp && t.reduced_cols == 1, "self-move keeps the model");
  expect(t.record.col_red2org == std::vector<int32_t>{0}, "self-move keeps the record");
  expect(host(t.reduced_lp.c, 1)[0] == 3.0, "self-move keeps device data");
  gp::GpuPresolveSummary borrowed;
  borrowed.reduced_lp = t.reduced_lp;  // not owned
  gp::GpuPresolveSummary borrowed2 = std::move(borrowed);
  expect(!borrowed2.owns_reduced_lp, "a borrowed model stays borrowed when moved");
  gp::free_gpu_presolve_reduced_lp(borrowed2);  // must not free t's arrays
  expect(borrowed2.reduced_lp.c == nullptr, "free clears a borrowed view");
  expect(host(t.reduced_lp.c, 1)[0] == 3.0, "borrowed free left the owner's data alone");
  gp::free_gpu_presolve_reduced_lp(t);
  gp::free_gpu_presolve_reduced_lp(t);  // idempotent
  expect(!t.owns_reduced_lp && t.reduced_lp.c == nullptr, "free releases and clears");
}

// ---------------------------------------------------------------------------
// Two rows over three columns: x0 + x1 <= 2, x0 + x1 - x2 <= 3, with
// 0 <= x <= (10, 1, inf).  The implied bounds of row 0 tighten x0 to 2 and
// then row 1 is redundant: one row, three columns and two entries remain.
gp::LPInfoGpu propagation_lp() {
  gp::LPInfoGpu lp;
  lp.A = {2, 3, 5, dev<int32_t>({0, 2, 5}), dev<int32_t>({0, 1, 0, 1, 2}), dev<double>({1, 1, 1, 1, -1})};
  lp.AT = {3, 2, 5, dev<int32_t>({0, 2, 4, 5}), dev<int32_t>({0, 1, 0, 1, 1}), dev<double>({1, 1, 1, 1, -1})};
  lp.c = dev<double>({0, 0, 0});
  lp.l = dev<double>({0, 0, 0});
  lp.u = dev<double>({10, 1, kInf});
  lp.AL = dev<double>({-kInf, -kInf});
  lp.AU = dev<double>({2, 3});
  return lp;
}

gp::PresolveParams propagation_params(bool tiered) {
  gp::PresolveParams p;
  p.max_iters = tiered ? 2 : 1;
  p.use_tiered_scheduler = tiered;
  p.enable_tiered_bootstrap = false;
  p.record_postsolve_tape = true;
  p.record_postsolve_tape_cpu = tiered;
  p.enable_infeasible_fixed_variables = false;
  p.enable_empty_rows = false;
  p.enable_singleton_rows = false;
  p.enable_duplicate_rows = false;
  p.enable_empty_cols = false;
  p.enable_column_singletons_eq = false;
  p.enable_column_singletons_dual_infer = false;
  p.enable_doubleton_equations = false;
  p.enable_dual_fix = false;
  p.enable_duplicate_columns = false;
  p.enable_structural_l1_substitution = false;
  p.enable_redundant_bounds = false;
  return p;
}

void presolve_postsolve_flow(bool tiered) {
  g_case = tiered ? "presolve_postsolve_tiered" : "presolve_postsolve_fixed";
  gp::LPInfoGpu lp = propagation_lp();
  const gp::PresolveParams prm = propagation_params(tiered);
  {
    const gp::GpuPresolveSummary s = gp::run_gpu_presolve_with_record(lp, prm);
    expect(!s.has_infeasible && !s.has_unbounded, "feasible and bounded");
    expect(s.reduced_rows == 1 && s.reduced_cols == 3 && s.reduced_nnz == 2, "reduced to 1 x 3 with 2 entries");
    expect(s.record.row_red2org == std::vector<int32_t>{0}, "row 0 kept");
    expect(s.record.row_org2red == std::vector<int32_t>({0, -1}), "row 1 removed");
    expect(s.record.tape.types.size() == static_cast<size_t>(tiered ? 2 : 1), "tape length");
    expect(s.record.tape_gpu.record_count == 0, "the device tape mirror stays empty");
    expect(!s.owns_reduced_lp, "with_record keeps no reduced model");
    const int32_t t0 = s.record.tape.types[0];
    expect(t0 == static_cast<int32_t>(gp::PostsolveReductionType::BoundChangeTheRow) ||
               t0 == static_cast<int32_t>(gp::PostsolveReductionType::BoundChangeNoRow),
           "first record is a bound change");
    double* xr = dev<double>({1, 1, 0});
    double* yr = dev<double>({0});
    double* zr = dev<double>({0, 0, 0});
    gp::GpuPostsolveResult post = gp::postsolve_gpu(xr, yr, zr, s.record, &lp);
    const std::vector<double> x = host(post.x_org, 3), y = host(post.y_org, 2), z = host(post.z_org, 3);
    expect(post.n0 == 3 && post.m0 == 2, "postsolve sizes");
    expect(std::fabs(x[0] - 1) < 1e-12 && std::fabs(x[1] - 1) < 1e-12 && std::fabs(x[2]) < 1e-12, "x");
    expect(std::fabs(y[0]) < 1e-12 && std::fabs(y[1]) < 1e-12, "y");
    expect(std::fabs(z[0]) < 1e-12 && std::fabs(z[1]) < 1e-12 && std::fabs(z[2]) < 1e-12, "z");
    for (double* p : {xr, yr, zr, post.x_org, post.y_org, post.z_org}) hip_ok(hipFree(p), "hipFree outputs");
    // Null reduced inputs are allowed (postsolve_gpu with x/y/z absent).
    gp::GpuPostsolveResult empty = gp::postsolve_gpu(nullptr, nullptr, nullptr, s.record);
    expect(empty.n0 == 3 && empty.m0 == 2, "postsolve of null inputs");
    for (double* p : {empty.x_org, empty.y_org, empty.z_org}) hip_ok(hipFree(p), "hipFree outputs");
  }
  {
    gp::GpuPresolveSummary s = gp::run_gpu_presolve_with_reduced_lp(lp, prm);
    expect(s.owns_reduced_lp, "with_reduced_lp owns the reduced model");
    expect(s.reduced_lp.A.rows == 1 && s.reduced_lp.A.cols == 3 && s.reduced_lp.A.nnz == 2, "reduced model shape");
    const std::vector<int32_t> ptr = host(s.reduced_lp.A.rowPtr, 2);
    expect(ptr[0] == 0 && ptr[1] == 2, "reduced row pointers");
    const std::vector<double> u = host(s.reduced_lp.u, 3);
    expect(u[0] == 2.0 && u[1] == 1.0 && std::isinf(u[2]), "tightened upper bound of x0");
    gp::GpuPresolveSummary moved = std::move(s);
    gp::free_gpu_presolve_reduced_lp(moved);
    gp::free_gpu_presolve_reduced_lp(moved);
    gp::free_gpu_presolve_reduced_lp(s);
  }
  free_lp(lp);
}

// Three rows, one fixed and one free column: the scheduler removes
// everything (cpp/tests/test_gpu_direct_port.cpp:281-330).
void scheduler_removes_everything(bool tiered) {
  g_case = tiered ? "scheduler_empties_lp_tiered" : "scheduler_empties_lp_fixed";
  gp::LPInfoGpu lp;
  lp.A = {3, 3, 2, dev<int32_t>({0, 1, 2, 2}), dev<int32_t>({0, 1}), dev<double>({1, 1})};
  lp.AT = {3, 3, 2, dev<int32_t>({0, 1, 2, 2}), dev<int32_t>({0, 1}), dev<double>({1, 1})};
  lp.c = dev<double>({0, 0, 2});
  lp.l = dev<double>({5, 0, 1});
  lp.u = dev<double>({5, 10, 3});
  lp.AL = dev<double>({5, -kInf, -1});
  lp.AU = dev<double>({5, kInf, 1});
  gp::PresolveParams prm;
  prm.max_iters = 4;
  prm.use_tiered_scheduler = tiered;
  const gp::GpuPresolveSummary s = gp::run_gpu_presolve_with_record(lp, prm);
  expect(!s.has_infeasible && !s.has_unbounded, "feasible and bounded");
  expect(s.reduced_rows == 0 && s.reduced_cols == 0, "everything removed");
  free_lp(lp);
}

// ---------------------------------------------------------------------------
void rules_on_partial_plans() {
  g_case = "empty_rows_partial_plan";
  {
    gp::LPInfoGpu lp;  // only the dimensions: the rule reads no CSR array
    lp.A.rows = 3;
    lp.A.cols = 2;
    gp::PresolvePlanGpu plan;  // no column vectors, no objective
    plan.keep_row_mask = dev<uint8_t>({1, 1, 1});
    plan.new_AL = dev<double>({-1, -2, 0});
    plan.new_AU = dev<double>({1, 3, 0});
    gp::PresolveStatsGpu st;
    st.row_nnz = dev<int32_t>({0, 0, 1});
    gp::PresolveParams prm;
    prm.feasibility_tol = 1e-9;
    gp::apply_rule_empty_rows(plan, lp, st, prm);
    const std::vector<uint8_t> keep = host(plan.keep_row_mask, 3);
    expect(keep[0] == 0 && keep[1] == 0 && keep[2] == 1, "empty feasible rows removed");
    expect(plan.has_change && plan.has_row_action && !plan.has_infeasible, "plan flags");
    for (void* p : {static_cast<void*>(plan.keep_row_mask), static_cast<void*>(plan.new_AL),
                    static_cast<void*>(plan.new_AU), static_cast<void*>(st.row_nnz)})
      (void)hipFree(p);
  }
  g_case = "empty_rows_infeasible";
  {
    gp::LPInfoGpu lp;
    lp.A.rows = 2;
    lp.A.cols = 1;
    gp::PresolvePlanGpu plan;
    plan.keep_row_mask = dev<uint8_t>({1, 1});
    plan.new_AL = dev<double>({-1, 2});
    plan.new_AU = dev<double>({1, 3});
    gp::PresolveStatsGpu st;
    st.row_nnz = dev<int32_t>({0, 0});
    gp::PresolveParams prm;
    prm.feasibility_tol = 1e-9;
    gp::apply_rule_empty_rows(plan, lp, st, prm);
    const std::vector<uint8_t> keep = host(plan.keep_row_mask, 2);
    expect(keep[0] == 1 && keep[1] == 1, "no removal once infeasible");
    expect(plan.has_infeasible, "infeasible empty row detected");
    for (void* p : {static_cast<void*>(plan.keep_row_mask), static_cast<void*>(plan.new_AL),
                    static_cast<void*>(plan.new_AU), static_cast<void*>(st.row_nnz)})
      (void)hipFree(p);
  }
  g_case = "singleton_rows_partial_plan";
  {
    // Row 0: 2 x0 in [2, 6] -> x0 in [1, 3]; row 1 keeps two entries.
    gp::LPInfoGpu lp;
    lp.A = {2, 2, 3, dev<int32_t>({0, 1, 3}), dev<int32_t>({0, 0, 1}), dev<double>({2, 1, 1})};
    lp.AT = {2, 2, 3, dev<int32_t>({0, 2, 3}), dev<int32_t>({0, 1, 1}), dev<double>({2, 1, 1})};
    gp::PresolvePlanGpu plan;  // no objective vector
    plan.keep_row_mask = dev<uint8_t>({1, 1});
    plan.keep_col_mask = dev<uint8_t>({1, 1});
    plan.new_l = dev<double>({0, 0});
    plan.new_u = dev<double>({10, 10});
    plan.new_AL = dev<double>({2, -kInf});
    plan.new_AU = dev<double>({6, 4});
    gp::PresolveStatsGpu st;
    st.row_nnz = dev<int32_t>({1, 2});
    st.singleton_row_mask = dev<uint8_t>({1, 0});
    st.singleton_row_col = dev<int32_t>({0, -1});
    st.singleton_row_val = dev<double>({2, 0});
    gp::PresolveParams prm;
    gp::apply_rule_singleton_rows(plan, lp, st, prm);
    const std::vector<uint8_t> keep = host(plan.keep_row_mask, 2);
    const std::vector<double> l = host(plan.new_l, 2), u = host(plan.new_u, 2);
    expect(keep[0] == 0 && keep[1] == 1, "singleton row removed");
    expect(l[0] == 1.0 && u[0] == 3.0 && l[1] == 0.0 && u[1] == 10.0, "bounds of x0 from the row");
    expect(plan.has_change && !plan.has_infeasible, "plan flags");
    for (void* p : {static_cast<void*>(plan.keep_row_mask), static_cast<void*>(plan.keep_col_mask),
                    static_cast<void*>(plan.new_l), static_cast<void*>(plan.new_u), static_cast<void*>(plan.new_AL),
                    static_cast<void*>(plan.new_AU), static_cast<void*>(st.row_nnz),
                    static_cast<void*>(st.singleton_row_mask), static_cast<void*>(st.singleton_row_col),
                    static_cast<void*>(st.singleton_row_val)})
      (void)hipFree(p);
    free_lp(lp);
  }
}

void progress_rule() {
  g_case = "has_good_nnz_progress";
  expect(!gp::_has_good_nnz_progress(0, 0, 0.95), "no progress from nothing");
  expect(gp::_has_good_nnz_progress(100, 94, 0.95), "94 < 95");
  expect(!gp::_has_good_nnz_progress(100, 95, 0.95), "95 is not below 95");
}

void errors_are_runtime_errors() {
  g_case = "input_errors";
  gp::LPInfoGpu lp = propagation_lp();
  std::vector<int32_t> bad{0, 0, 1, 0, 2};  // row 0 holds column 0 twice
  hip_ok(hipMemcpy(lp.A.colVal, bad.data(), sizeof(int32_t) * 5, hipMemcpyHostToDevice), "hipMemcpy");
  bool threw = false;
  try {
    (void)gp::run_gpu_presolve_with_record(lp, gp::PresolveParams{});
  } catch (const std::runtime_error& e) {
    threw = std::string(e.what()).find("repeats a column") != std::string::npos;
  }
  expect(threw, "a repeated column index throws std::runtime_error");
  // The thread's next call works normally.
  std::vector<int32_t> good{0, 1, 0, 1, 2};
  hip_ok(hipMemcpy(lp.A.colVal, good.data(), sizeof(int32_t) * 5, hipMemcpyHostToDevice), "hipMemcpy");
  const gp::GpuPresolveSummary s = gp::run_gpu_presolve_with_record(lp, propagation_params(false));
  expect(s.reduced_rows == 1 && s.reduced_nnz == 2, "next call after an error");
  free_lp(lp);
}

}  // namespace

int main() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
    std::printf("FAIL no HIP device\n");
    return 1;
  }
  struct Case {
    const char* name;
    void (*fn)();
  };
  const Case cases[] = {
      {"summary_ownership", summary_ownership},
      {"presolve_postsolve_fixed", [] { presolve_postsolve_flow(false); }},
      {"presolve_postsolve_tiered", [] { presolve_postsolve_flow(true); }},
      {"scheduler_empties_lp_tiered", [] { scheduler_removes_everything(true); }},
      {"scheduler_empties_lp_fixed", [] { scheduler_removes_everything(false); }},
      {"rules_on_partial_plans", rules_on_partial_plans},
      {"has_good_nnz_progress", progress_rule},
      {"input_errors", errors_are_runtime_errors},
  };
  for (const Case& c : cases) {
    const int before = g_failures;
    try {
      c.fn();
    } catch (const std::exception& e) {
      std::printf("FAIL %s: exception: %s\n", c.name, e.what());
      ++g_failures;
    }
    if (g_failures == before) std::printf("ok %s\n", c.name);
    std::fflush(stdout);
  }
  return g_failures == 0 ? 0 : 1;
}
