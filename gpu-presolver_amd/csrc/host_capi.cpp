// C-ABI of libgpu_presolver_amd.so (include/gp_presolve_capi.h): flat
// wrappers over the C++ API, exceptions turned into gp_status codes with the
// message kept per thread for gp_last_error().
#include "gp_presolve_capi.h"

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <memory>
#include <sstream>
#include <new>
#include <stdexcept>
#include <thread>
#include <string>
#include <vector>

#include "gp_host.hpp"
#include "gpu_presolver/presolve/gpu_postsolve.hpp"
#include "gpu_presolver/presolve/gpu_presolve.hpp"
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
#include "gpu_presolver/tools.hpp"
#include "mps_reader/mpsreader.hpp"

using namespace gpu_presolver::presolve;
namespace amd = gpu_presolver::amd;

struct gp_presolve_handle {
  GpuPresolveSummary summary;
  bool keep_reduced = false;
};

struct gp_tape_handle {
  PostsolveTape tape;
};

struct gp_server {
  amd::Server* impl = nullptr;
};

// A read MPS model with int32 copies of its index arrays for the view.
struct gp_mps_model {
  mps_reader::MpsModel mps;
  std::vector<int32_t> csr_ptr, csr_idx, csc_ptr, csc_idx;
  void fill() {
    const auto& a = mps.lp.csr();
    const auto& t = mps.lp.csc();
    csr_ptr.assign(a.row_ptr().begin(), a.row_ptr().end());
    csr_idx.assign(a.col_idx().begin(), a.col_idx().end());
    csc_ptr.assign(t.col_ptr().begin(), t.col_ptr().end());
    csc_idx.assign(t.row_idx().begin(), t.row_idx().end());
  }
};

namespace {

thread_local std::string t_error;

struct InvalidArgument : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};

template <class F>
gp_status guarded(F&& f) {
  try {
    amd::AbandonOnThrow abandon_guard;
    f();
    t_error.clear();
    return GP_OK;
  } catch (const InvalidArgument& e) {
    t_error = e.what();
    return GP_ERR_INVALID_ARGUMENT;
  } catch (const amd::InputError& e) {
    t_error = e.what();
    return GP_ERR_INVALID_ARGUMENT;
  } catch (const amd::UnsupportedStructure& e) {
    t_error = e.what();
    return GP_ERR_UNSUPPORTED;
  } catch (const amd::HipError& e) {
    t_error = e.what();
    return e.status() == hipErrorNoDevice || e.status() == hipErrorInvalidDevice ? GP_ERR_NO_DEVICE : GP_ERR_DEVICE;
  } catch (const std::bad_alloc&) {
    t_error = "host allocation failed";
    return GP_ERR_INTERNAL;
  } catch (const std::exception& e) {
    t_error = e.what();
    return GP_ERR_INTERNAL;
  } catch (...) {
    t_error = "unknown exception";
    return GP_ERR_INTERNAL;
  }
}

void require(bool cond, const char* what) {
  if (!cond) throw InvalidArgument(what);
}

void require_device() {
  int count = 0;
  const hipError_t st = hipGetDeviceCount(&count);
  if (st != hipSuccess || count <= 0) throw amd::HipError("no HIP device", hipErrorNoDevice);
}

DeviceCsrMatrix to_csr(const gp_csr& a) { return DeviceCsrMatrix{a.rows, a.cols, a.nnz, a.row_ptr, a.col_idx, a.val}; }
gp_csr from_csr(const DeviceCsrMatrix& a) { return gp_csr{a.rows, a.cols, a.nnz, a.rowPtr, a.colVal, a.nzVal}; }

void check_csr(const gp_csr& a, const char* name) {
  if (a.rows < 0 || a.cols < 0 || a.nnz < 0) throw InvalidArgument(std::string(name) + ": negative dimension");
  if (a.row_ptr == nullptr) throw InvalidArgument(std::string(name) + ": null row_ptr");
  if (a.nnz > 0 && (a.col_idx == nullptr || a.val == nullptr))
    throw InvalidArgument(std::string(name) + ": null col_idx/val");
}

LPInfoGpu to_lp(const gp_lp& p) {
  check_csr(p.A, "lp.A");
  check_csr(p.AT, "lp.AT");
  require(p.AT.rows == p.A.cols && p.AT.cols == p.A.rows && p.AT.nnz == p.A.nnz, "lp.AT is not shaped as A^T");
  LPInfoGpu lp;
  lp.A = to_csr(p.A);
  lp.AT = to_csr(p.AT);
  lp.c = p.c;
  lp.AL = p.AL;
  lp.AU = p.AU;
  lp.l = p.l;
  lp.u = p.u;
  lp.obj_constant = p.obj_constant;
  return lp;
}

gp_lp from_lp(const LPInfoGpu& lp) {
  gp_lp p;
  p.A = from_csr(lp.A);
  p.AT = from_csr(lp.AT);
  p.c = lp.c;
  p.AL = lp.AL;
  p.AU = lp.AU;
  p.l = lp.l;
  p.u = lp.u;
  p.obj_constant = lp.obj_constant;
  return p;
}

gpk_lpgen_spec to_spec(const gp_lp_spec& q) {
  return gpk_lpgen_spec{q.seed, q.m, q.n, q.len_lo, q.len_hi, q.mean_len, q.min_len};
}

PresolveParams to_params(const gp_params* q) {
  PresolveParams p;
  if (q == nullptr) return p;
  p.max_iters = q->max_iters;
  p.max_time = q->max_time;
  p.verbose = q->verbose != 0;
  p.record_postsolve_tape = q->record_postsolve_tape != 0;
  p.record_postsolve_tape_cpu = q->record_postsolve_tape_cpu != 0;
  p.feasibility_tol = q->feasibility_tol;
  p.bound_tol = q->bound_tol;
  p.zero_tol = q->zero_tol;
  p.implied_variable_bounds_max_rounds = q->implied_variable_bounds_max_rounds;
  p.implied_variable_bounds_max_bound_only_rounds = q->implied_variable_bounds_max_bound_only_rounds;
  p.implied_variable_bounds_preserve_free_zero_cost = q->implied_variable_bounds_preserve_free_zero_cost != 0;
  p.doubleton_equations_single_batch_per_iter = q->doubleton_equations_single_batch_per_iter != 0;
  p.doubleton_equations_max_fill_in_proxy = q->doubleton_equations_max_fill_in_proxy;
  p.doubleton_equations_scan = q->doubleton_equations_scan != 0;
  p.doubleton_equations_min_selected_per_batch = q->doubleton_equations_min_selected_per_batch;
  p.doubleton_equations_min_selected_ratio = q->doubleton_equations_min_selected_ratio;
  p.doubleton_equations_max_batch_rounds = q->doubleton_equations_max_batch_rounds;
  p.doubleton_equations_max_time = q->doubleton_equations_max_time;
  p.enable_infeasible_fixed_variables = q->enable_infeasible_fixed_variables != 0;
  p.enable_empty_rows = q->enable_empty_rows != 0;
  p.enable_singleton_rows = q->enable_singleton_rows != 0;
  p.enable_infeasible_redundant_rows = q->enable_infeasible_redundant_rows != 0;
  p.enable_implied_variable_bounds = q->enable_implied_variable_bounds != 0;
  p.enable_duplicate_rows = q->enable_duplicate_rows != 0;
  p.enable_empty_cols = q->enable_empty_cols != 0;
  p.enable_column_singletons_eq = q->enable_column_singletons_eq != 0;
  p.enable_column_singletons_dual_infer = q->enable_column_singletons_dual_infer != 0;
  p.enable_doubleton_equations = q->enable_doubleton_equations != 0;
  p.enable_dual_fix = q->enable_dual_fix != 0;
  p.dual_fix_max_nonzero_cost_col_degree = q->dual_fix_max_nonzero_cost_col_degree;
  p.enable_duplicate_columns = q->enable_duplicate_columns != 0;
  p.duplicate_columns_max_abs_ratio = q->duplicate_columns_max_abs_ratio;
  p.duplicate_columns_scale_guard_max_group_size = q->duplicate_columns_scale_guard_max_group_size;
  p.duplicate_columns_scale_guard_min_groups = q->duplicate_columns_scale_guard_min_groups;
  p.enable_structure_specific_rules = q->enable_structure_specific_rules != 0;
  p.enable_zero_cost_redundant_box_bounds = q->enable_zero_cost_redundant_box_bounds != 0;
  p.enable_redundant_bounds = q->enable_redundant_bounds != 0;
  p.use_tiered_scheduler = q->use_tiered_scheduler != 0;
  p.enable_tiered_bootstrap = q->enable_tiered_bootstrap != 0;
  p.implied_variable_bounds_bound_only_nnz_round_budget = q->implied_variable_bounds_bound_only_nnz_round_budget;
  p.enable_bounded_two_row_projection = q->enable_bounded_two_row_projection != 0;
  p.bounded_two_row_projection_max_row_nnz = q->bounded_two_row_projection_max_row_nnz;
  p.bounded_two_row_projection_safety_tol = q->bounded_two_row_projection_safety_tol;
  p.bounded_two_row_projection_min_candidates = q->bounded_two_row_projection_min_candidates;
  p.bounded_two_row_projection_min_candidate_ratio = q->bounded_two_row_projection_min_candidate_ratio;
  p.enable_orphan_mccormick_projection = q->enable_orphan_mccormick_projection != 0;
  p.orphan_mccormick_projection_min_candidates = q->orphan_mccormick_projection_min_candidates;
  p.orphan_mccormick_projection_min_candidate_ratio = q->orphan_mccormick_projection_min_candidate_ratio;
  p.enable_linf_components = q->enable_linf_components != 0;
  p.linf_probe_pairs = q->linf_probe_pairs;
  p.linf_min_pairs = q->linf_min_pairs;
  p.linf_max_tail_cols = q->linf_max_tail_cols;
  p.enable_structural_l1_substitution = q->enable_structural_l1_substitution != 0;
  p.enable_antipodal_components = q->enable_antipodal_components != 0;
  p.antipodal_probe_pairs = q->antipodal_probe_pairs;
  p.antipodal_min_pairs = q->antipodal_min_pairs;
  p.enable_covering_cost_dominance = q->enable_covering_cost_dominance != 0;
  p.covering_cost_dominance_min_fixed_cols = q->covering_cost_dominance_min_fixed_cols;
  p.covering_cost_dominance_min_candidate_ratio = q->covering_cost_dominance_min_candidate_ratio;
  return p;
}

void fill_tape_view(const PostsolveTape& t, gp_record_view* out) {
  out->tape_count = static_cast<int32_t>(t.types.size());
  out->tape_types = t.types.data();
  out->tape_index_starts = t.index_starts.data();
  out->tape_value_starts = t.value_starts.data();
  out->tape_dual_modes = t.dual_modes.data();
  out->tape_index_count = static_cast<int32_t>(t.indices.size());
  out->tape_indices = t.indices.data();
  out->tape_value_count = static_cast<int32_t>(t.vals.size());
  out->tape_vals = t.vals.data();
}

template <class T>
std::vector<T> host_vec(const T* p, int32_t n, const char* what) {
  if (n < 0) throw InvalidArgument(std::string(what) + ": negative length");
  if (n > 0 && p == nullptr) throw InvalidArgument(std::string(what) + ": null pointer");
  return n > 0 ? std::vector<T>(p, p + n) : std::vector<T>{};
}

PresolveRecordGpu to_record(const gp_record_view& v) {
  require(v.m0 >= 0 && v.n0 >= 0 && v.m1 >= 0 && v.n1 >= 0 && v.m1 <= v.m0 && v.n1 <= v.n0,
          "record: invalid dimensions");
  PresolveRecordGpu r;
  r.m0 = v.m0;
  r.n0 = v.n0;
  r.m1 = v.m1;
  r.n1 = v.n1;
  r.row_red2org = host_vec(v.row_red2org, v.m1, "record.row_red2org");
  r.col_red2org = host_vec(v.col_red2org, v.n1, "record.col_red2org");
  r.fixed_idx = host_vec(v.fixed_idx, v.fixed_count, "record.fixed_idx");
  r.fixed_val = host_vec(v.fixed_val, v.fixed_count, "record.fixed_val");
  r.obj_constant_old = v.obj_constant_old;
  r.obj_constant_new = v.obj_constant_new;
  r.tape.types = host_vec(v.tape_types, v.tape_count, "record.tape_types");
  r.tape.index_starts = host_vec(v.tape_index_starts, v.tape_count + 1, "record.tape_index_starts");
  r.tape.value_starts = host_vec(v.tape_value_starts, v.tape_count + 1, "record.tape_value_starts");
  r.tape.dual_modes = v.tape_dual_modes ? host_vec(v.tape_dual_modes, v.tape_count, "record.tape_dual_modes")
                                        : std::vector<uint8_t>(static_cast<size_t>(v.tape_count), 0);
  r.tape.indices = host_vec(v.tape_indices, v.tape_index_count, "record.tape_indices");
  r.tape.vals = host_vec(v.tape_vals, v.tape_value_count, "record.tape_vals");
  return r;
}

}  // namespace

extern "C" {

const char* gp_last_error(void) { return t_error.c_str(); }

const char* gp_version(void) { return "gpu_presolver_amd 0.1.0 (gfx950)"; }

void gp_params_default(gp_params* q) {
  if (q == nullptr) return;
  const PresolveParams p;
  q->max_iters = p.max_iters;
  q->max_time = p.max_time;
  q->verbose = p.verbose;
  q->record_postsolve_tape = p.record_postsolve_tape;
  q->record_postsolve_tape_cpu = p.record_postsolve_tape_cpu;
  q->feasibility_tol = p.feasibility_tol;
  q->bound_tol = p.bound_tol;
  q->zero_tol = p.zero_tol;
  q->implied_variable_bounds_max_rounds = p.implied_variable_bounds_max_rounds;
  q->implied_variable_bounds_max_bound_only_rounds = p.implied_variable_bounds_max_bound_only_rounds;
  q->implied_variable_bounds_preserve_free_zero_cost = p.implied_variable_bounds_preserve_free_zero_cost;
  q->doubleton_equations_single_batch_per_iter = p.doubleton_equations_single_batch_per_iter;
  q->doubleton_equations_max_fill_in_proxy = p.doubleton_equations_max_fill_in_proxy;
  q->doubleton_equations_scan = p.doubleton_equations_scan;
  q->doubleton_equations_min_selected_per_batch = p.doubleton_equations_min_selected_per_batch;
  q->doubleton_equations_min_selected_ratio = p.doubleton_equations_min_selected_ratio;
  q->doubleton_equations_max_batch_rounds = p.doubleton_equations_max_batch_rounds;
  q->doubleton_equations_max_time = p.doubleton_equations_max_time;
  q->enable_infeasible_fixed_variables = p.enable_infeasible_fixed_variables;
  q->enable_empty_rows = p.enable_empty_rows;
  q->enable_singleton_rows = p.enable_singleton_rows;
  q->enable_infeasible_redundant_rows = p.enable_infeasible_redundant_rows;
  q->enable_implied_variable_bounds = p.enable_implied_variable_bounds;
  q->enable_duplicate_rows = p.enable_duplicate_rows;
  q->enable_empty_cols = p.enable_empty_cols;
  q->enable_column_singletons_eq = p.enable_column_singletons_eq;
  q->enable_column_singletons_dual_infer = p.enable_column_singletons_dual_infer;
  q->enable_doubleton_equations = p.enable_doubleton_equations;
  q->enable_dual_fix = p.enable_dual_fix;
  q->dual_fix_max_nonzero_cost_col_degree = p.dual_fix_max_nonzero_cost_col_degree;
  q->enable_duplicate_columns = p.enable_duplicate_columns;
  q->duplicate_columns_max_abs_ratio = p.duplicate_columns_max_abs_ratio;
  q->duplicate_columns_scale_guard_max_group_size = p.duplicate_columns_scale_guard_max_group_size;
  q->duplicate_columns_scale_guard_min_groups = p.duplicate_columns_scale_guard_min_groups;
  q->enable_structure_specific_rules = p.enable_structure_specific_rules;
  q->enable_zero_cost_redundant_box_bounds = p.enable_zero_cost_redundant_box_bounds;
  q->enable_redundant_bounds = p.enable_redundant_bounds;
  q->use_tiered_scheduler = p.use_tiered_scheduler;
  q->enable_tiered_bootstrap = p.enable_tiered_bootstrap;
  q->implied_variable_bounds_bound_only_nnz_round_budget = p.implied_variable_bounds_bound_only_nnz_round_budget;
  q->enable_bounded_two_row_projection = p.enable_bounded_two_row_projection;
  q->bounded_two_row_projection_max_row_nnz = p.bounded_two_row_projection_max_row_nnz;
  q->bounded_two_row_projection_safety_tol = p.bounded_two_row_projection_safety_tol;
  q->bounded_two_row_projection_min_candidates = p.bounded_two_row_projection_min_candidates;
  q->bounded_two_row_projection_min_candidate_ratio = p.bounded_two_row_projection_min_candidate_ratio;
  q->enable_orphan_mccormick_projection = p.enable_orphan_mccormick_projection;
  q->orphan_mccormick_projection_min_candidates = p.orphan_mccormick_projection_min_candidates;
  q->orphan_mccormick_projection_min_candidate_ratio = p.orphan_mccormick_projection_min_candidate_ratio;
  q->enable_linf_components = p.enable_linf_components;
  q->linf_probe_pairs = p.linf_probe_pairs;
  q->linf_min_pairs = p.linf_min_pairs;
  q->linf_max_tail_cols = p.linf_max_tail_cols;
  q->enable_structural_l1_substitution = p.enable_structural_l1_substitution;
  q->enable_antipodal_components = p.enable_antipodal_components;
  q->antipodal_probe_pairs = p.antipodal_probe_pairs;
  q->antipodal_min_pairs = p.antipodal_min_pairs;
  q->enable_covering_cost_dominance = p.enable_covering_cost_dominance;
  q->covering_cost_dominance_min_fixed_cols = p.covering_cost_dominance_min_fixed_cols;
  q->covering_cost_dominance_min_candidate_ratio = p.covering_cost_dominance_min_candidate_ratio;
}

gp_status gp_device_free(void* device_ptr) {
  return guarded([&] {
    if (device_ptr == nullptr) return;
    amd::dev_free(device_ptr, amd::ctx().stream());
    amd::ctx().sync("gp_device_free");
  });
}

gp_status gp_row_activity(const gp_csr* A, const double* l, const double* u, double zero_tol, double* row_min_fin,
                          double* row_max_fin, int32_t* row_min_neg_inf_count, int32_t* row_max_pos_inf_count,
                          void* stream) {
  return guarded([&] {
    require(A != nullptr, "A is null");
    check_csr(*A, "A");
    require(A->rows == 0 || (row_min_fin && row_max_fin && row_min_neg_inf_count && row_max_pos_inf_count),
            "null output");
    require(A->cols == 0 || (l && u), "null bounds");
    if (stream != nullptr) {
      if (A->rows == 0) return;
      amd::row_activity(A->row_ptr, A->col_idx, A->val, A->rows, A->cols, l, u, zero_tol, row_min_fin, row_max_fin,
                        row_min_neg_inf_count, row_max_pos_inf_count, static_cast<hipStream_t>(stream));
      return;
    }
    require_device();
    amd::check(hipDeviceSynchronize(), "entry synchronize");
    compute_row_activity_summary(row_min_fin, row_max_fin, row_min_neg_inf_count, row_max_pos_inf_count, to_csr(*A),
                                 l, u, zero_tol);
    amd::ctx().sync("gp_row_activity");
  });
}

gp_status gp_row_activity_noalloc(const gp_csr* A, const double* l, const double* u, double zero_tol,
                                  double* row_min_fin, double* row_max_fin, int32_t* row_min_neg_inf_count,
                                  int32_t* row_max_pos_inf_count, void* stream) {
  // No library context, pool or staging area is touched here: nothing in
  // this call allocates.
  return guarded([&] {
    require(A != nullptr, "A is null");
    check_csr(*A, "A");
    require(A->rows == 0 || (row_min_fin && row_max_fin && row_min_neg_inf_count && row_max_pos_inf_count),
            "null output");
    require(A->cols == 0 || (l && u), "null bounds");
    require(A->cols <= GP_ROW_ACTIVITY_NOALLOC_MAX_COLS,
            "gp_row_activity_noalloc: more columns than its fixed staging array holds (2,097,152)");
    if (A->rows == 0) return;
    require_device();
    hipStream_t s = static_cast<hipStream_t>(stream);
    amd::check(gpk_row_activity_noalloc(A->row_ptr, A->col_idx, A->val, A->rows, A->cols, l, u, zero_tol,
                                        row_min_fin, row_max_fin, row_min_neg_inf_count, row_max_pos_inf_count, s),
               "gp_row_activity_noalloc");
    if (s == nullptr) amd::check(hipStreamSynchronize(nullptr), "gp_row_activity_noalloc synchronize");
  });
}

gp_status gp_col_max_abs(const gp_csr* AT, double* col_max_abs, void* stream) {
  return guarded([&] {
    require(AT != nullptr, "AT is null");
    check_csr(*AT, "AT");
    require(AT->rows == 0 || col_max_abs, "null output");
    if (stream != nullptr) {
      if (AT->rows == 0) return;
      amd::check(gpk_col_max_abs(AT->row_ptr, AT->val, AT->rows, col_max_abs, static_cast<hipStream_t>(stream)),
                 "col max abs");
      return;
    }
    require_device();
    amd::check(hipDeviceSynchronize(), "entry synchronize");
    compute_col_max_abs(col_max_abs, to_csr(*AT));
    amd::ctx().sync("gp_col_max_abs");
  });
}

gp_status gp_compact_csr(const gp_csr* A, const uint8_t* keep_row, const uint8_t* keep_col, gp_csr* out) {
  return guarded([&] {
    require(A != nullptr && out != nullptr, "null argument");
    check_csr(*A, "A");
    require((A->rows == 0 || keep_row) && (A->cols == 0 || keep_col), "null keep mask");
    require_device();
    amd::check(hipDeviceSynchronize(), "entry synchronize");
    amd::DBuf<int32_t> row_map, col_map;
    const int32_t m_new = amd::build_old_to_new(keep_row, A->rows, row_map);
    const int32_t n_new = amd::build_old_to_new(keep_col, A->cols, col_map);
    const DeviceCsrMatrix r = amd::compact_csr(to_csr(*A), row_map, col_map, m_new, n_new);
    amd::ctx().sync("gp_compact_csr");
    *out = from_csr(r);
  });
}

gp_status gp_transpose_csr(const gp_csr* A, gp_csr* out) {
  return guarded([&] {
    require(A != nullptr && out != nullptr, "null argument");
    check_csr(*A, "A");
    require_device();
    amd::check(hipDeviceSynchronize(), "entry synchronize");
    const DeviceCsrMatrix r = amd::transpose_csr(to_csr(*A));
    amd::ctx().sync("gp_transpose_csr");
    *out = from_csr(r);
  });
}

gp_status gp_rule_apply(int32_t rule_id, gp_plan_view* plan, const gp_lp* lp, const gp_stats_view* stats,
                        const gp_params* params, const uint8_t* candidate_col_mask, gp_tape_handle** tape) {
  return guarded([&] {
    require(plan != nullptr && lp != nullptr, "null plan or lp");
    require(rule_id >= GP_RULE_EMPTY_ROWS && rule_id <= GP_RULE_REDUNDANT_BOUNDS, "unknown rule id");
    const LPInfoGpu L = to_lp(*lp);
    require_device();
    amd::check(hipDeviceSynchronize(), "entry synchronize");
    PresolvePlanGpu P;
    P.keep_row_mask = plan->keep_row;
    P.keep_col_mask = plan->keep_col;
    P.new_c = plan->c;
    P.new_l = plan->l;
    P.new_u = plan->u;
    P.new_AL = plan->AL;
    P.new_AU = plan->AU;
    P.obj_constant_delta = plan->obj_constant_delta;
    P.has_change = plan->has_change != 0;
    P.has_row_action = plan->has_row_action != 0;
    P.has_col_action = plan->has_col_action != 0;
    P.has_infeasible = plan->has_infeasible != 0;
    P.has_unbounded = plan->has_unbounded != 0;
    P.has_new_A = plan->has_new_A != 0;
    if (P.has_new_A) {
      check_csr(plan->new_A, "plan.new_A");
      P.new_A = to_csr(plan->new_A);
    }
    PresolveStatsGpu S;
    if (stats != nullptr) {
      S.row_nnz = stats->row_nnz;
      S.singleton_row_mask = stats->singleton_row_mask;
      S.singleton_row_col = stats->singleton_row_col;
      S.singleton_row_val = stats->singleton_row_val;
      S.col_nnz = stats->col_nnz;
      S.empty_col_mask = stats->empty_col_mask;
      S.column_singleton_mask = stats->column_singleton_mask;
      S.column_singleton_row = stats->column_singleton_row;
      S.column_singleton_val = stats->column_singleton_val;
    }
    const PresolveParams prm = to_params(params);
    switch (rule_id) {
      case GP_RULE_EMPTY_ROWS: apply_rule_empty_rows(P, L, S, prm); break;
      case GP_RULE_SINGLETON_ROWS: apply_rule_singleton_rows(P, L, S, prm); break;
      case GP_RULE_INFEASIBLE_REDUNDANT_ROWS: apply_rule_infeasible_redundant_rows(P, L, S, prm); break;
      case GP_RULE_IMPLIED_VARIABLE_BOUNDS: apply_rule_implied_variable_bounds(P, L, S, prm); break;
      case GP_RULE_DUPLICATE_ROWS: apply_rule_duplicate_rows(P, L, S, prm); break;
      case GP_RULE_INFEASIBLE_FIXED_VARIABLES: apply_rule_infeasible_fixed_variables(P, L, prm); break;
      case GP_RULE_EMPTY_COLS: apply_rule_empty_cols(P, L, S, prm); break;
      case GP_RULE_COLUMN_SINGLETONS_DUAL_INFER: apply_rule_column_singletons_dual_infer(P, L, S, prm); break;
      case GP_RULE_COLUMN_SINGLETONS_EQ: apply_rule_column_singletons_eq(P, L, S, prm, false); break;
      case GP_RULE_DOUBLETON_EQUATIONS: apply_rule_doubleton_equations(P, L, S, prm); break;
      case GP_RULE_DUAL_FIX: apply_rule_dual_fix(P, L, S, prm, candidate_col_mask); break;
      case GP_RULE_DUPLICATE_COLUMNS: apply_rule_duplicate_columns(P, L, S, prm); break;
      case GP_RULE_REDUNDANT_BOUNDS: apply_rule_redundant_bounds(P, L, S, prm); break;
      default: break;
    }
    amd::ctx().sync("gp_rule_apply");
    plan->obj_constant_delta = P.obj_constant_delta;
    plan->has_change = P.has_change;
    plan->has_row_action = P.has_row_action;
    plan->has_col_action = P.has_col_action;
    plan->has_infeasible = P.has_infeasible;
    plan->has_unbounded = P.has_unbounded;
    plan->has_new_A = P.has_new_A;
    if (P.has_new_A) plan->new_A = from_csr(P.new_A);
    if (tape != nullptr && !P.tape.types.empty()) {
      if (*tape == nullptr) *tape = new gp_tape_handle();
      PostsolveTape& dst = (*tape)->tape;
      const int32_t bi = static_cast<int32_t>(dst.indices.size()), bv = static_cast<int32_t>(dst.vals.size());
      dst.types.insert(dst.types.end(), P.tape.types.begin(), P.tape.types.end());
      dst.dual_modes.insert(dst.dual_modes.end(), P.tape.dual_modes.begin(), P.tape.dual_modes.end());
      dst.indices.insert(dst.indices.end(), P.tape.indices.begin(), P.tape.indices.end());
      dst.vals.insert(dst.vals.end(), P.tape.vals.begin(), P.tape.vals.end());
      for (size_t k = 1; k < P.tape.index_starts.size(); ++k) {
        dst.index_starts.push_back(bi + P.tape.index_starts[k]);
        dst.value_starts.push_back(bv + P.tape.value_starts[k]);
      }
    }
  });
}

gp_status gp_tape_view(const gp_tape_handle* tape, gp_record_view* out) {
  return guarded([&] {
    require(out != nullptr, "null output");
    *out = gp_record_view{};
    static const PostsolveTape kEmpty;
    fill_tape_view(tape ? tape->tape : kEmpty, out);
  });
}

void gp_tape_free(gp_tape_handle* tape) { delete tape; }

gp_status gp_presolve(const gp_lp* lp, const gp_params* params, int32_t keep_reduced_lp, gp_presolve_handle** out) {
  return guarded([&] {
    require(lp != nullptr && out != nullptr, "null argument");
    *out = nullptr;
    const LPInfoGpu L = to_lp(*lp);
    require_device();
    const PresolveParams prm = to_params(params);
    auto* h = new gp_presolve_handle();
    try {
      h->keep_reduced = keep_reduced_lp != 0;
      h->summary = h->keep_reduced ? run_gpu_presolve_with_reduced_lp(L, prm) : run_gpu_presolve_with_record(L, prm);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

gp_status gp_presolve_summary(const gp_presolve_handle* h, gp_summary* out) {
  return guarded([&] {
    require(h != nullptr && out != nullptr, "null argument");
    const GpuPresolveSummary& s = h->summary;
    out->original_rows = s.original_rows;
    out->original_cols = s.original_cols;
    out->reduced_rows = s.reduced_rows;
    out->reduced_cols = s.reduced_cols;
    out->reduced_nnz = s.reduced_nnz;
    out->iterations = s.iterations;
    out->elapsed_seconds = s.elapsed_seconds;
    out->obj_constant_delta = s.obj_constant_delta;
    out->has_infeasible = s.has_infeasible;
    out->has_unbounded = s.has_unbounded;
  });
}

gp_status gp_presolve_record(const gp_presolve_handle* h, gp_record_view* out) {
  return guarded([&] {
    require(h != nullptr && out != nullptr, "null argument");
    const PresolveRecordGpu& r = h->summary.record;
    *out = gp_record_view{};
    out->m0 = r.m0;
    out->n0 = r.n0;
    out->m1 = r.m1;
    out->n1 = r.n1;
    out->row_red2org = r.row_red2org.data();
    out->col_red2org = r.col_red2org.data();
    out->fixed_count = static_cast<int32_t>(r.fixed_idx.size());
    out->fixed_idx = r.fixed_idx.data();
    out->fixed_val = r.fixed_val.data();
    out->obj_constant_old = r.obj_constant_old;
    out->obj_constant_new = r.obj_constant_new;
    fill_tape_view(r.tape, out);
  });
}

gp_status gp_presolve_reduced_lp(const gp_presolve_handle* h, gp_lp* out) {
  return guarded([&] {
    require(h != nullptr && out != nullptr, "null argument");
    require(h->keep_reduced, "presolve ran without keep_reduced_lp");
    *out = from_lp(h->summary.reduced_lp);
  });
}

int32_t gp_has_good_nnz_progress(int32_t nnz_before, int32_t nnz_after, double ratio) {
  return _has_good_nnz_progress(nnz_before, nnz_after, ratio) ? 1 : 0;
}

gp_status gp_server_create(const gp_params* params, const gp_serve_config* config, gp_server** out) {
  return guarded([&] {
    require(config != nullptr && out != nullptr, "null argument");
    require(config->workers >= 1, "gp_serve_config.workers must be >= 1");
    *out = nullptr;
    require_device();
    auto* h = new gp_server();
    try {
      h->impl = amd::server_create(to_params(params), *config);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

gp_status gp_server_submit(gp_server* server, const gp_lp* lp, uint64_t tag, double t_arrival, int64_t* job) {
  return guarded([&] {
    require(server != nullptr && lp != nullptr, "null argument");
    const LPInfoGpu L = to_lp(*lp);
    const int64_t id = amd::server_submit(server->impl, L, tag, t_arrival);
    if (job) *job = id;
  });
}

int32_t gp_server_wait(gp_server* server, gp_job_result* out, int32_t max, double timeout_s) {
  if (server == nullptr || out == nullptr || max <= 0) return 0;
  return amd::server_wait(server->impl, out, max, timeout_s);
}

double gp_server_now(const gp_server* server) { return server ? amd::server_now(server->impl) : 0.0; }

void gp_server_destroy(gp_server* server) {
  if (server == nullptr) return;
  amd::server_destroy(server->impl);
  delete server;
}

gp_status gp_serve_trace(const gp_lp* lps, int32_t lp_count, const int32_t* lp_of_job, const double* arrival_s,
                         int32_t jobs, const gp_params* params, const gp_serve_config* config,
                         gp_job_result* results) {
  return guarded([&] {
    require(lps != nullptr && lp_of_job != nullptr && config != nullptr && results != nullptr, "null argument");
    require(lp_count > 0 && jobs >= 0 && config->workers >= 1, "invalid trace size or worker count");
    std::vector<LPInfoGpu> L;
    for (int32_t k = 0; k < lp_count; ++k) L.push_back(to_lp(lps[k]));
    for (int32_t k = 0; k < jobs; ++k) {
      require(lp_of_job[k] >= 0 && lp_of_job[k] < lp_count, "lp_of_job out of range");
      if (arrival_s) require(k == 0 || arrival_s[k] >= arrival_s[k - 1], "arrival times must be non-decreasing");
    }
    require_device();
    std::unique_ptr<amd::Server, void (*)(amd::Server*)> srv(amd::server_create(to_params(params), *config),
                                                             amd::server_destroy);
    for (int32_t k = 0; k < jobs; ++k) {
      const double t = arrival_s ? arrival_s[k] : 0.0;
      const double wait = t - amd::server_now(srv.get());
      if (wait > 0.0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      amd::server_submit(srv.get(), L[lp_of_job[k]], static_cast<uint64_t>(k), t);
    }
    std::vector<gp_job_result> buf(64);
    for (int32_t got = 0; got < jobs;) {
      const int n = amd::server_wait(srv.get(), buf.data(), static_cast<int>(buf.size()), 1.0);
      for (int i = 0; i < n; ++i) results[buf[i].tag] = buf[i];
      got += n;
    }
  });
}

gp_status gp_plain_job(const gp_lp* lp, const gp_params* params, uint64_t seed, uint64_t tag, int32_t postsolve,
                       gp_job_result* out) {
  return guarded([&] {
    require(lp != nullptr && out != nullptr, "null argument");
    const LPInfoGpu L = to_lp(*lp);
    require_device();
    const auto clock = [] {
      return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    };
    const double t0 = clock();
    *out = amd::run_job(L, to_params(params), postsolve != 0, seed, tag);
    out->job = 0;
    out->worker = -1;
    out->t_arrival = out->t_start = out->t_ready = t0;
    out->t_end = clock();
  });
}

gp_status gp_plain_job_spec(const gp_lp_spec* spec, const gp_params* params, uint64_t seed, uint64_t tag,
                            int32_t postsolve, gp_job_result* out) {
  return guarded([&] {
    require(spec != nullptr && out != nullptr, "null argument");
    const gpk_lpgen_spec sp = to_spec(*spec);
    amd::check_spec(sp);
    require_device();
    const auto clock = [] {
      return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    };
    const double t0 = clock();
    double t_ready = t0;
    *out = amd::run_generated_job(sp, to_params(params), postsolve != 0, seed, tag, &t_ready);
    out->job = 0;
    out->worker = -1;
    out->t_arrival = out->t_start = t0;
    out->t_ready = t_ready;
    out->t_end = clock();
  });
}

gp_status gp_generate_lp(const gp_lp_spec* spec, gp_lp* out) {
  return guarded([&] {
    require(spec != nullptr && out != nullptr, "null argument");
    const gpk_lpgen_spec sp = to_spec(*spec);
    amd::check_spec(sp);
    require_device();
    LPInfoGpu lp = amd::generate_lp(sp);
    try {
      amd::ctx().sync("gp_generate_lp");
    } catch (...) {
      amd::free_generated_lp(lp);
      throw;
    }
    *out = from_lp(lp);
  });
}

gp_status gp_free_lp(gp_lp* lp) {
  return guarded([&] {
    if (lp == nullptr) return;
    LPInfoGpu L;
    L.A = to_csr(lp->A);
    L.AT = to_csr(lp->AT);
    L.c = lp->c;
    L.l = lp->l;
    L.u = lp->u;
    L.AL = lp->AL;
    L.AU = lp->AU;
    amd::free_generated_lp(L);
    amd::ctx().sync("gp_free_lp");
    *lp = gp_lp{};
  });
}

gp_status gp_server_submit_spec(gp_server* server, const gp_lp_spec* spec, uint64_t tag, double t_arrival,
                                int64_t* job) {
  return guarded([&] {
    require(server != nullptr && spec != nullptr, "null argument");
    const int64_t id = amd::server_submit_spec(server->impl, to_spec(*spec), tag, t_arrival);
    if (job) *job = id;
  });
}

gp_status gp_serve_spec_trace(const gp_lp_spec* specs, int32_t jobs, const double* arrival_s, const gp_params* params,
                              const gp_serve_config* config, gp_job_result* results) {
  return guarded([&] {
    require(specs != nullptr && config != nullptr && results != nullptr, "null argument");
    require(jobs >= 0 && config->workers >= 1, "invalid trace size or worker count");
    for (int32_t k = 0; k < jobs; ++k) {
      amd::check_spec(to_spec(specs[k]));
      if (arrival_s) require(k == 0 || arrival_s[k] >= arrival_s[k - 1], "arrival times must be non-decreasing");
    }
    require_device();
    std::unique_ptr<amd::Server, void (*)(amd::Server*)> srv(amd::server_create(to_params(params), *config),
                                                             amd::server_destroy);
    for (int32_t k = 0; k < jobs; ++k) {
      const double t = arrival_s ? arrival_s[k] : 0.0;
      const double wait = t - amd::server_now(srv.get());
      if (wait > 0.0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      amd::server_submit_spec(srv.get(), to_spec(specs[k]), static_cast<uint64_t>(k), t);
    }
    std::vector<gp_job_result> buf(64);
    for (int32_t got = 0; got < jobs;) {
      const int n = amd::server_wait(srv.get(), buf.data(), static_cast<int>(buf.size()), 1.0);
      for (int i = 0; i < n; ++i) results[buf[i].tag] = buf[i];
      got += n;
    }
  });
}

gp_status gp_synthetic_point(uint64_t seed, const gp_lp* lp, double* x, double* y, double* z) {
  return guarded([&] {
    require(lp != nullptr, "null argument");
    const LPInfoGpu L = to_lp(*lp);
    require((L.A.cols == 0 || (x && z)) && (L.A.rows == 0 || y), "null output vector");
    require_device();
    amd::Ctx& c = amd::ctx();
    amd::check(gpk_synthetic_point(seed, L.A.cols, L.A.rows, L.l, L.u, x, y, z, c.stream()), "synthetic point");
    c.sync("synthetic point");
  });
}

gp_status gp_checksum(const double* x, int32_t n, const double* y, int32_t m, const double* z, uint64_t* out) {
  return guarded([&] {
    require(out != nullptr && n >= 0 && m >= 0, "invalid argument");
    require((n == 0 || (x && z)) && (m == 0 || y), "null vector");
    require_device();
    *out = amd::checksum_xyz(x, n, y, m, z);
  });
}

void gp_set_batched_closure_min_nnz(int32_t nnz) { amd::set_batched_closure_min_nnz(nnz); }
int32_t gp_batched_closure_min_nnz(void) { return amd::batched_closure_min_nnz(); }
void gp_test_fail_after_record_updates(int32_t count) { amd::fail_after_record_updates(count); }
void gp_test_check_plan_reuse(int32_t on) { amd::check_plan_reuse(on != 0); }
int64_t gp_test_plan_reuses(void) { return amd::plan_reuses(); }

gp_status gp_mps_read(const char* path, gp_mps_model** out) {
  return guarded([&] {
    require(path != nullptr && out != nullptr, "null argument");
    *out = nullptr;
    auto m = std::make_unique<gp_mps_model>(gp_mps_model{gpu_presolver::tools::read_mps_file(path), {}, {}, {}, {}});
    m->fill();
    *out = m.release();
  });
}

gp_status gp_mps_parse(const char* text, size_t length, gp_mps_model** out) {
  return guarded([&] {
    require((text != nullptr || length == 0) && out != nullptr, "null argument");
    *out = nullptr;
    std::istringstream in(std::string(text ? text : "", length));
    auto m = std::make_unique<gp_mps_model>(gp_mps_model{mps_reader::read_mps(in), {}, {}, {}, {}});
    m->fill();
    *out = m.release();
  });
}

gp_status gp_mps_get_view(const gp_mps_model* model, gp_mps_view* out) {
  return guarded([&] {
    require(model != nullptr && out != nullptr, "null argument");
    const mps_reader::LpModel& lp = model->mps.lp;
    out->name = model->mps.name.c_str();
    out->rows = lp.num_rows();
    out->cols = lp.num_cols();
    out->nnz = lp.csr().nnz();
    out->csr_row_ptr = model->csr_ptr.data();
    out->csr_col_idx = model->csr_idx.data();
    out->csr_val = lp.csr().values().data();
    out->csc_col_ptr = model->csc_ptr.data();
    out->csc_row_idx = model->csc_idx.data();
    out->csc_val = lp.csc().values().data();
    out->c = lp.objective().data();
    out->row_lower = lp.row_lower().data();
    out->row_upper = lp.row_upper().data();
    out->col_lower = lp.col_lower().data();
    out->col_upper = lp.col_upper().data();
    out->obj_constant = lp.obj_constant();
  });
}

gp_status gp_mps_upload(const gp_mps_model* model, gp_lp* out) {
  return guarded([&] {
    require(model != nullptr && out != nullptr, "null argument");
    require_device();
    gpu_presolver::tools::DeviceLpOwner owner = gpu_presolver::tools::upload_lp(model->mps.lp);
    *out = from_lp(owner.lp);
    owner.lp = LPInfoGpu{};  // the caller owns the arrays now (gp_free_lp)
  });
}

void gp_mps_free(gp_mps_model* model) { delete model; }

void gp_presolve_free(gp_presolve_handle* h) {
  if (h == nullptr) return;
  free_gpu_presolve_reduced_lp(h->summary);
  delete h;
}

gp_status gp_postsolve(const gp_record_view* record, const double* x_red, const double* y_red, const double* z_red,
                       const gp_lp* original, double* x_org, double* y_org, double* z_org) {
  return guarded([&] {
    require(record != nullptr, "null record");
    const PresolveRecordGpu rec = to_record(*record);
    require(rec.n0 == 0 || (x_org && z_org), "null x_org/z_org");
    require(rec.m0 == 0 || y_org, "null y_org");
    LPInfoGpu orig;
    if (original != nullptr) orig = to_lp(*original);
    require_device();
    amd::PostsolvePlan::validate(rec, original ? &orig : nullptr);
    if (amd::job_mode().device_sync_on_entry) amd::check(hipDeviceSynchronize(), "gp_postsolve entry synchronize");
    // Straight into the caller's outputs.
    amd::PostsolvePlan plan(rec, original ? &orig : nullptr);
    plan.enqueue(x_red, y_red, z_red, x_org, y_org, z_org);
    amd::ctx().sync("gp_postsolve");
  });
}

}  // extern "C"

struct gp_postsolve_plan {
  PresolveRecordGpu record;
  LPInfoGpu orig;
  bool has_orig = false;
  bool use_graph = false;
  bool graph_failed = false;
  std::unique_ptr<amd::PostsolvePlan> plan;
  amd::DBuf<double> in_x, in_y, in_z, out_x, out_y, out_z;  // the graph's own inputs and outputs
  hipGraphExec_t exec = nullptr;
  int64_t graph_runs = 0;
  ~gp_postsolve_plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

namespace {

// Captures the plan's launch sequence over its own buffers; false (and the
// stream left out of capture) when the runtime refuses.
bool capture_postsolve(gp_postsolve_plan& p) {
  const amd::PostsolvePlan& plan = *p.plan;
  p.in_x.alloc(plan.n1());
  p.in_y.alloc(plan.m1());
  p.in_z.alloc(plan.n1());
  p.out_x.alloc(plan.n0());
  p.out_y.alloc(plan.m0());
  p.out_z.alloc(plan.n0());
  hipStream_t s = amd::ctx().stream();
  amd::ctx().sync("postsolve graph capture");  // nothing of this stream's earlier work enters the graph
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  bool ok = true;
  try {
    p.plan->enqueue(p.in_x, p.in_y, p.in_z, p.out_x, p.out_y, p.out_z);
  } catch (...) {
    ok = false;
  }
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(s, &g);
  if (!ok || end != hipSuccess || g == nullptr) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return false;
  }
  const hipError_t inst = hipGraphInstantiate(&p.exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (inst != hipSuccess) {
    p.exec = nullptr;
    (void)hipGetLastError();
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

gp_status gp_postsolve_prepare(const gp_record_view* record, const gp_lp* original, int32_t use_graph,
                               gp_postsolve_plan** out) {
  return guarded([&] {
    require(record != nullptr && out != nullptr, "null record or output handle");
    *out = nullptr;
    auto p = std::make_unique<gp_postsolve_plan>();
    p->record = to_record(*record);
    if (original != nullptr) {
      p->orig = to_lp(*original);
      p->has_orig = true;
    }
    p->use_graph = use_graph != 0;
    require_device();
    amd::PostsolvePlan::validate(p->record, p->has_orig ? &p->orig : nullptr);
    if (amd::job_mode().device_sync_on_entry) amd::check(hipDeviceSynchronize(), "gp_postsolve_prepare entry");
    p->plan = std::make_unique<amd::PostsolvePlan>(p->record, p->has_orig ? &p->orig : nullptr);
    amd::ctx().sync("gp_postsolve_prepare");
    *out = p.release();
  });
}

gp_status gp_postsolve_plan_run(gp_postsolve_plan* p, const double* x_red, const double* y_red, const double* z_red,
                                double* x_org, double* y_org, double* z_org) {
  return guarded([&] {
    require(p != nullptr && p->plan, "null postsolve plan");
    amd::PostsolvePlan& plan = *p->plan;
    require(plan.n0() == 0 || (x_org && z_org), "null x_org/z_org");
    require(plan.m0() == 0 || y_org, "null y_org");
    require_device();
    if (amd::job_mode().device_sync_on_entry) amd::check(hipDeviceSynchronize(), "gp_postsolve_plan_run entry");
    const bool all_inputs = x_red && (y_red || plan.m1() == 0) && z_red;
    if (p->use_graph && all_inputs && !p->graph_failed && p->exec == nullptr && !capture_postsolve(*p))
      p->graph_failed = true;
    if (p->use_graph && all_inputs && p->exec != nullptr) {
      hipStream_t s = amd::ctx().stream();
      amd::CopyBatch()
          .copy(p->in_x.get(), x_red, plan.n1())
          .copy(p->in_y.get(), y_red, plan.m1())
          .copy(p->in_z.get(), z_red, plan.n1())
          .run("postsolve graph inputs");
      amd::check(hipGraphLaunch(p->exec, s), "postsolve graph launch");
      amd::CopyBatch()
          .copy(x_org, p->out_x.get(), plan.n0())
          .copy(y_org, p->out_y.get(), plan.m0())
          .copy(z_org, p->out_z.get(), plan.n0())
          .run("postsolve graph outputs");
      ++p->graph_runs;
    } else {
      plan.enqueue(x_red, y_red, z_red, x_org, y_org, z_org);
    }
    amd::ctx().sync("gp_postsolve_plan_run");
  });
}

int64_t gp_postsolve_plan_graph_runs(const gp_postsolve_plan* p) { return p ? p->graph_runs : 0; }

gp_status gp_postsolve_plan_shape(const gp_postsolve_plan* p, gp_postsolve_shape* out) {
  return guarded([&] {
    require(p != nullptr && p->plan != nullptr && out != nullptr, "null argument");
    const amd::PostsolvePlan::Shape sh = p->plan->shape();
    *out = gp_postsolve_shape{sh.records, sh.primal_levels, sh.dual_levels, sh.tape_bytes};
  });
}

void gp_postsolve_plan_free(gp_postsolve_plan* p) {
  if (p == nullptr) return;
  (void)guarded([&] { amd::ctx().sync("gp_postsolve_plan_free"); });
  delete p;
}

}  // extern "C"

extern "C" {

void gp_pdhg_params_default(gp_pdhg_params* params) {
  if (params == nullptr) return;
  const amd::PdhgOptions d;
  *params = gp_pdhg_params{d.max_iters, d.max_time, d.tol, d.check_every, d.ruiz_iters, d.power_iters, d.use_graph,
                           d.verbose, d.method};
}

namespace {
amd::PdhgOptions pdhg_options(const gp_pdhg_params* params) {
  amd::PdhgOptions o;
  if (params != nullptr)
    o = amd::PdhgOptions{params->max_iters,   params->max_time,    params->tol,
                         params->check_every, params->ruiz_iters,  params->power_iters,
                         params->use_graph,   params->verbose,     params->method};
  return o;
}
gp_pdhg_result pdhg_result(const amd::PdhgReport& r) {
  return gp_pdhg_result{r.status,   r.iterations,      r.restarts,      r.checks,        r.primal_obj,
                        r.dual_obj, r.primal_res,      r.dual_res,      r.gap,           r.elapsed_seconds,
                        r.norm_estimate, r.primal_weight, r.graph_launches};
}
}  // namespace

gp_status gp_presolve_solve(const gp_lp* lp, const gp_params* params, const gp_pdhg_params* solve_params,
                            double* x, double* y, double* z, gp_pipeline_result* out) {
  return guarded([&] {
    require(lp != nullptr && out != nullptr, "null argument");
    *out = gp_pipeline_result{};
    const LPInfoGpu L = to_lp(*lp);
    require(L.A.cols == 0 || (x && z), "null x/z");
    require(L.A.rows == 0 || y, "null y");
    require_device();
    const PresolveParams prm = to_params(params);
    const amd::PipelineReport r = amd::presolve_solve(L, prm, pdhg_options(solve_params), x, y, z);
    out->reduced_rows = r.reduced_rows;
    out->reduced_cols = r.reduced_cols;
    out->reduced_nnz = r.reduced_nnz;
    out->presolve_iterations = r.iterations;
    out->obj_constant_delta = r.obj_constant_delta;
    out->has_infeasible = r.has_infeasible ? 1 : 0;
    out->has_unbounded = r.has_unbounded ? 1 : 0;
    out->solve = pdhg_result(r.solve);
    out->presolve_seconds = r.presolve_seconds;
    out->solve_seconds = r.solve_seconds;
    out->prepare_seconds = r.prepare_seconds;
    out->postsolve_seconds = r.postsolve_seconds;
    out->total_seconds = r.total_seconds;
  });
}

gp_status gp_pdhg_solve(const gp_lp* lp, const gp_pdhg_params* params, double* x, double* y, double* z,
                        gp_pdhg_result* out) {
  return guarded([&] {
    require(lp != nullptr && out != nullptr, "null argument");
    *out = gp_pdhg_result{};
    const LPInfoGpu L = to_lp(*lp);
    require(L.A.cols == 0 || (L.c && L.l && L.u), "null c/l/u");
    require(L.A.rows == 0 || (L.AL && L.AU), "null AL/AU");
    require_device();
    const amd::PdhgReport r = amd::pdhg_solve(L, pdhg_options(params), x, y, z);
    amd::ctx().sync("gp_pdhg_solve");
    *out = pdhg_result(r);
  });
}

}  // extern "C"
