// Tool layer (include/gpu_presolver/tools.hpp): environment overrides, the
// device LP upload, binary vector files, the reduced-LP export directory and
// the postsolve checker's hashes and feasibility figures.  Semantics of the
// reference's tool header and tools (cpp/tools/gpu_presolver_tool_common.hpp,
// gpu_presolver_export_reduced.cpp:16-57,
// gpu_presolver_check_postsolve_solution.cpp:20-246).
#include "gpu_presolver/tools.hpp"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <memory>

namespace gpu_presolver::tools {

namespace {

std::string lower(const char* s) {
  std::string t(s);
  for (char& ch : t) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
  return t;
}

template <class T>
void env_bool(const char* name, T& field) {
  field = parse_bool_env_value(std::getenv(name), field);
}

std::vector<std::int32_t> as_i32(const std::vector<int>& v) { return std::vector<std::int32_t>(v.begin(), v.end()); }

}  // namespace

bool parse_bool_env_value(const char* value, bool fallback) {
  if (value == nullptr) return fallback;
  const std::string t = lower(value);
  if (t == "1" || t == "true" || t == "yes" || t == "on") return true;
  if (t == "0" || t == "false" || t == "no" || t == "off") return false;
  return fallback;
}

std::int64_t parse_int64_env_value(const char* value, std::int64_t fallback) {
  if (value == nullptr) return fallback;
  char* end = nullptr;
  errno = 0;
  const long long v = std::strtoll(value, &end, 10);
  if (errno != 0 || end == value || *end != '\0') return fallback;
  return static_cast<std::int64_t>(v);
}

double parse_double_env_value(const char* value, double fallback) {
  if (value == nullptr) return fallback;
  char* end = nullptr;
  errno = 0;
  const double v = std::strtod(value, &end);
  if (errno != 0 || end == value || *end != '\0' || !std::isfinite(v)) return fallback;
  return v;
}

void apply_presolve_env_overrides(presolve::PresolveParams& p) {
  if (const char* v = std::getenv("GPUPRESOLVER_MAX_ITERS")) p.max_iters = std::max(0, std::atoi(v));
  if (const char* v = std::getenv("GPUPRESOLVER_SCHEDULER")) {
    const std::string t = lower(v);
    if (t == "fixed" || t == "legacy" || t == "old") p.use_tiered_scheduler = false;
    else if (t == "tiered" || t == "default") p.use_tiered_scheduler = true;
  }
  env_bool("GPUPRESOLVER_USE_TIERED_SCHEDULER", p.use_tiered_scheduler);
  env_bool("GPUPRESOLVER_ENABLE_DOUBLETON_EQUATIONS", p.enable_doubleton_equations);
  env_bool("GPUPRESOLVER_ENABLE_DUAL_FIX", p.enable_dual_fix);
  env_bool("GPUPRESOLVER_ENABLE_DUPLICATE_COLUMNS", p.enable_duplicate_columns);
  p.dual_fix_max_nonzero_cost_col_degree = static_cast<int>(parse_int64_env_value(
      std::getenv("GPUPRESOLVER_DUAL_FIX_MAX_NONZERO_COST_COL_DEGREE"), p.dual_fix_max_nonzero_cost_col_degree));
  p.duplicate_columns_max_abs_ratio = parse_double_env_value(
      std::getenv("GPUPRESOLVER_DUPLICATE_COLUMNS_MAX_ABS_RATIO"), p.duplicate_columns_max_abs_ratio);
  env_bool("GPUPRESOLVER_IMPLIED_VARIABLE_BOUNDS_PRESERVE_FREE_ZERO_COST",
           p.implied_variable_bounds_preserve_free_zero_cost);
  p.implied_variable_bounds_bound_only_nnz_round_budget =
      parse_int64_env_value(std::getenv("GPUPRESOLVER_IMPLIED_VARIABLE_BOUNDS_BOUND_ONLY_NNZ_ROUND_BUDGET"),
                            p.implied_variable_bounds_bound_only_nnz_round_budget);
  env_bool("GPUPRESOLVER_ENABLE_REDUNDANT_BOUNDS", p.enable_redundant_bounds);
  env_bool("GPUPRESOLVER_ENABLE_ZERO_COST_REDUNDANT_BOX_BOUNDS", p.enable_zero_cost_redundant_box_bounds);
  env_bool("GPUPRESOLVER_ENABLE_STRUCTURE_SPECIFIC_RULES", p.enable_structure_specific_rules);
  env_bool("GPUPRESOLVER_ENABLE_STRUCTURAL_L1_SUBSTITUTION", p.enable_structural_l1_substitution);
  env_bool("GPUPRESOLVER_ENABLE_LINF_COMPONENTS", p.enable_linf_components);
  env_bool("GPUPRESOLVER_ENABLE_ANTIPODAL_COMPONENTS", p.enable_antipodal_components);
  env_bool("GPUPRESOLVER_ENABLE_COVERING_COST_DOMINANCE", p.enable_covering_cost_dominance);
  env_bool("GPUPRESOLVER_ENABLE_BOUNDED_TWO_ROW_PROJECTION", p.enable_bounded_two_row_projection);
  env_bool("GPUPRESOLVER_ENABLE_ORPHAN_MCCORMICK_PROJECTION", p.enable_orphan_mccormick_projection);
}

void check_hip(hipError_t status, const char* context) {
  if (status != hipSuccess) throw std::runtime_error(std::string(context) + ": " + hipGetErrorString(status));
}

void DeviceLpOwner::release() noexcept {
  for (void* p : {static_cast<void*>(lp.A.rowPtr), static_cast<void*>(lp.A.colVal), static_cast<void*>(lp.A.nzVal),
                  static_cast<void*>(lp.AT.rowPtr), static_cast<void*>(lp.AT.colVal), static_cast<void*>(lp.AT.nzVal),
                  static_cast<void*>(lp.c), static_cast<void*>(lp.AL), static_cast<void*>(lp.AU),
                  static_cast<void*>(lp.l), static_cast<void*>(lp.u)})
    if (p) (void)hipFree(p);
  lp = {};
}

DeviceLpOwner upload_lp(const mps_reader::LpModel& model) {
  DeviceLpOwner o;
  const mps_reader::CsrMatrix& a = model.csr();
  const mps_reader::CscMatrix& t = model.csc();
  o.lp.A.rows = a.rows();
  o.lp.A.cols = a.cols();
  o.lp.A.nnz = a.nnz();
  o.lp.A.rowPtr = copy_to_device(as_i32(a.row_ptr()));
  o.lp.A.colVal = copy_to_device(as_i32(a.col_idx()));
  o.lp.A.nzVal = copy_to_device(a.values());
  o.lp.AT.rows = t.cols();
  o.lp.AT.cols = t.rows();
  o.lp.AT.nnz = t.nnz();
  o.lp.AT.rowPtr = copy_to_device(as_i32(t.col_ptr()));
  o.lp.AT.colVal = copy_to_device(as_i32(t.row_idx()));
  o.lp.AT.nzVal = copy_to_device(t.values());
  o.lp.c = copy_to_device(model.objective());
  o.lp.AL = copy_to_device(model.row_lower());
  o.lp.AU = copy_to_device(model.row_upper());
  o.lp.l = copy_to_device(model.col_lower());
  o.lp.u = copy_to_device(model.col_upper());
  o.lp.obj_constant = model.obj_constant();
  return o;
}

mps_reader::MpsModel read_mps_file(const std::string& path) {
  if (path == "-") return mps_reader::read_mps(std::cin);
  std::ifstream in(path);
  if (!in.is_open()) throw std::runtime_error("failed to open " + path);
  return mps_reader::read_mps(in);
}

void close_output_file(std::ofstream& out, const std::filesystem::path& path) {
  out.flush();
  const bool flushed = static_cast<bool>(out);
  out.close();
  if (!flushed || !out) throw std::runtime_error("failed to write " + path.string());
}

void write_lp_export(const std::filesystem::path& dir, const std::string& model_name,
                     const presolve::GpuPresolveSummary& s) {
  const presolve::LPInfoGpu& lp = s.reduced_lp;
  std::filesystem::create_directories(dir);
  std::ofstream meta(dir / "meta.txt");
  if (!meta.is_open()) throw std::runtime_error("failed to open export meta");
  meta << std::setprecision(std::numeric_limits<double>::max_digits10);
  const char* status = s.has_infeasible ? "infeasible" : (s.has_unbounded ? "unbounded" : "ok");
  meta << "format gpu_presolver_reduced_lp_v1\n"
       << "model " << model_name << "\n"
       << "status " << status << "\n"
       << "original_rows " << s.original_rows << "\n"
       << "original_cols " << s.original_cols << "\n"
       << "reduced_rows " << s.reduced_rows << "\n"
       << "reduced_cols " << s.reduced_cols << "\n"
       << "reduced_nnz " << s.reduced_nnz << "\n"
       << "obj_constant " << lp.obj_constant << "\n"
       << "objective_shift " << s.obj_constant_delta << "\n"
       << "iterations " << s.iterations << "\n";
  close_output_file(meta, dir / "meta.txt");
  const auto m = static_cast<std::size_t>(lp.A.rows), n = static_cast<std::size_t>(lp.A.cols),
             nnz = static_cast<std::size_t>(lp.A.nnz);
  write_binary_vector(dir / "A_rowPtr_i32.bin", copy_to_host(lp.A.rowPtr, m + 1));
  write_binary_vector(dir / "A_colVal_i32.bin", copy_to_host(lp.A.colVal, nnz));
  write_binary_vector(dir / "A_nzVal_f64.bin", copy_to_host(lp.A.nzVal, nnz));
  write_binary_vector(dir / "c_f64.bin", copy_to_host(lp.c, n));
  write_binary_vector(dir / "l_f64.bin", copy_to_host(lp.l, n));
  write_binary_vector(dir / "u_f64.bin", copy_to_host(lp.u, n));
  write_binary_vector(dir / "AL_f64.bin", copy_to_host(lp.AL, m));
  write_binary_vector(dir / "AU_f64.bin", copy_to_host(lp.AU, m));
}

void Fnv1a64::add_bytes(const void* data, std::size_t size) {
  const auto* b = static_cast<const unsigned char*>(data);
  for (std::size_t i = 0; i < size; ++i) {
    hash_ ^= static_cast<std::uint64_t>(b[i]);
    hash_ *= 1099511628211ull;
  }
}

std::uint64_t record_hash(const presolve::PresolveRecordGpu& r) {
  Fnv1a64 h;
  h.add_scalar(r.m0);
  h.add_scalar(r.n0);
  h.add_scalar(r.m1);
  h.add_scalar(r.n1);
  h.add_vector(r.row_org2red);
  h.add_vector(r.row_red2org);
  h.add_vector(r.col_org2red);
  h.add_vector(r.col_red2org);
  h.add_vector(r.fixed_idx);
  h.add_vector(r.fixed_val);
  h.add_vector(r.removed_row_idx);
  h.add_vector(r.removed_col_idx);
  h.add_scalar(r.obj_constant_old);
  h.add_scalar(r.obj_constant_new);
  h.add_scalar(static_cast<std::uint8_t>(r.has_primal_only_antipodal_reduction));
  h.add_scalar(static_cast<std::uint8_t>(r.has_primal_only_covering_cost_dominance_reduction));
  h.add_scalar(static_cast<std::uint8_t>(r.has_primal_only_projected_auxiliary_reduction));
  h.add_scalar(static_cast<std::uint64_t>(r.structural_primal_recoveries.size()));
  for (const auto& s : r.structural_primal_recoveries) {
    h.add_string(s.pattern);
    h.add_scalar(static_cast<std::uint64_t>(s.splits.size()));
    for (const auto& x : s.splits) {
      h.add_scalar(x.t_col);
      h.add_scalar(x.e_col);
      h.add_scalar(x.rho);
    }
    h.add_scalar(static_cast<std::uint64_t>(s.outer_pairs.size()));
    for (const auto& x : s.outer_pairs) {
      h.add_scalar(x.bound_col);
      h.add_scalar(x.free_col);
    }
    h.add_scalar(static_cast<std::uint64_t>(s.linked_slacks.size()));
    for (const auto& x : s.linked_slacks) {
      h.add_scalar(x.slack_col);
      h.add_scalar(x.t_col);
      h.add_scalar(x.factor);
    }
    h.add_scalar(static_cast<std::uint64_t>(s.max_slacks.size()));
    for (const auto& x : s.max_slacks) {
      h.add_scalar(x.slack_col);
      h.add_vector(x.t_cols);
      h.add_vector(x.factors);
    }
  }
  h.add_scalar(static_cast<std::uint64_t>(r.antipodal_component_recoveries.size()));
  for (const auto& a : r.antipodal_component_recoveries) {
    h.add_vector(a.plus_cols);
    h.add_vector(a.minus_cols);
    h.add_vector(a.root_plus_cols);
    h.add_vector(a.root_minus_cols);
    h.add_vector(a.plus_lower);
    h.add_vector(a.minus_lower);
  }
  h.add_scalar(static_cast<std::uint64_t>(r.primal_recovery_timeline.size()));
  for (const auto& c : r.primal_recovery_timeline) {
    h.add_scalar(static_cast<std::uint8_t>(c.kind));
    h.add_scalar(c.payload_index);
    h.add_scalar(c.tape_position);
  }
  h.add_vector(r.tape.types);
  h.add_vector(r.tape.index_starts);
  h.add_vector(r.tape.value_starts);
  h.add_vector(r.tape.dual_modes);
  h.add_vector(r.tape.indices);
  h.add_vector(r.tape.vals);
  return h.value();
}

PrimalFeasibility compute_primal_feasibility(const mps_reader::LpModel& model, const std::vector<double>& x) {
  const mps_reader::CsrMatrix& A = model.csr();
  const std::vector<double>&AL = model.row_lower(), &AU = model.row_upper();
  const std::vector<double>&l = model.col_lower(), &u = model.col_upper();
  for (std::size_t j = 0; j < x.size(); ++j)
    if (!std::isfinite(x[j]))
      throw std::runtime_error("non-finite original primal solution at index " + std::to_string(j));
  PrimalFeasibility f;
  for (int i = 0; i < A.rows(); ++i) {
    double act = 0.0;
    for (int p = A.row_ptr()[i]; p < A.row_ptr()[i + 1]; ++p) act += A.values()[p] * x[A.col_idx()[p]];
    if (!std::isfinite(act)) throw std::runtime_error("non-finite row activity at index " + std::to_string(i));
    const double v = std::max({0.0, AL[i] - act, act - AU[i]});
    f.row_l2 = std::hypot(f.row_l2, v);
    f.row_max = std::max(f.row_max, v);
  }
  for (std::size_t j = 0; j < x.size(); ++j) {
    const double v = std::max({0.0, l[j] - x[j], x[j] - u[j]});
    f.col_l2 = std::hypot(f.col_l2, v);
    f.col_max = std::max(f.col_max, v);
  }
  double norm = 0.0;
  for (const std::vector<double>* side : {&AL, &AU})
    for (double b : *side)
      if (std::isfinite(b)) norm = std::hypot(norm, b);
  if (!std::isfinite(norm)) throw std::runtime_error("non-finite constraint bound norm");
  f.normalized = std::max(f.row_l2, f.col_l2) / (1.0 + norm);
  return f;
}

}  // namespace gpu_presolver::tools
