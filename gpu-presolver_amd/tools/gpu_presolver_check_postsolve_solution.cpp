// gpu_presolver_check_postsolve_solution <model.mps|-> <solution_dir>:
// presolve the model on the GPU, read a reduced-space solution
// (x_f64.bin, y_f64.bin, z_f64.bin of the reduced dimensions), postsolve it
// and report the original-space primal feasibility and FNV-1a hashes of the
// record and of x, y, z (reference
// cpp/tools/gpu_presolver_check_postsolve_solution.cpp).  Exit status 0 when
// the normalized primal violation is at most 1e-6, 1 when it is larger or
// presolve reports the LP infeasible or unbounded, 2 on an error (missing,
// short, long or non-finite solution vectors included).
// GPUPRESOLVER_PRINT_TOP_VIOLATIONS lists the ten largest row and column
// violations.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <iostream>
#include <string>
#include <vector>

#include "gpu_presolver/presolve/gpu_postsolve.hpp"
#include "gpu_presolver/presolve/gpu_presolve.hpp"
#include "gpu_presolver/tools.hpp"

namespace {

namespace tools = gpu_presolver::tools;

void require_finite(const std::vector<double>& v, const char* what) {
  for (std::size_t i = 0; i < v.size(); ++i)
    if (!std::isfinite(v[i])) throw std::runtime_error(std::string("non-finite ") + what + " at index " + std::to_string(i));
}

struct Violation {
  int index;
  double amount, value, lower, upper;
};

// The `limit` largest violations, largest first (ties by index).
void top(std::vector<Violation>& list, const Violation& v, std::size_t limit) {
  if (v.amount <= 0.0 || limit == 0) return;
  list.push_back(v);
  std::sort(list.begin(), list.end(), [](const Violation& a, const Violation& b) {
    return a.amount != b.amount ? a.amount > b.amount : a.index < b.index;
  });
  if (list.size() > limit) list.resize(limit);
}

void print_top_violations(const mps_reader::LpModel& lp, const std::vector<double>& x, std::size_t limit) {
  const mps_reader::CsrMatrix& A = lp.csr();
  std::vector<Violation> rows, cols;
  for (int i = 0; i < A.rows(); ++i) {
    double act = 0.0;
    for (int p = A.row_ptr()[i]; p < A.row_ptr()[i + 1]; ++p) act += A.values()[p] * x[A.col_idx()[p]];
    const double lo = lp.row_lower()[i], up = lp.row_upper()[i];
    top(rows, Violation{i, std::max({0.0, lo - act, act - up}), act, lo, up}, limit);
  }
  for (std::size_t j = 0; j < x.size(); ++j) {
    const double lo = lp.col_lower()[j], up = lp.col_upper()[j];
    top(cols, Violation{static_cast<int>(j), std::max({0.0, lo - x[j], x[j] - up}), x[j], lo, up}, limit);
  }
  for (std::size_t k = 0; k < rows.size(); ++k)
    std::cout << "top_row_violation_" << (k + 1) << " " << rows[k].index << "," << rows[k].amount << ","
              << rows[k].value << "," << rows[k].lower << "," << rows[k].upper << "\n";
  for (std::size_t k = 0; k < cols.size(); ++k)
    std::cout << "top_col_violation_" << (k + 1) << " " << cols[k].index << "," << cols[k].amount << ","
              << cols[k].value << "," << cols[k].lower << "," << cols[k].upper << "\n";
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: gpu_presolver_check_postsolve_solution <model.mps|-> <solution_dir>\n";
    return 2;
  }
  namespace ps = gpu_presolver::presolve;
  try {
    const std::filesystem::path dir = argv[2];
    mps_reader::MpsModel model = tools::read_mps_file(argv[1]);
    tools::DeviceLpOwner dev = tools::upload_lp(model.lp);
    ps::PresolveParams params;
    tools::apply_presolve_env_overrides(params);
    ps::GpuPresolveSummary s = ps::run_gpu_presolve_with_record(dev.lp, params);
    if (s.has_infeasible || s.has_unbounded) {
      std::cerr << "cannot check a solution: presolve reports " << (s.has_infeasible ? "infeasible" : "unbounded")
                << "\n";
      return 1;
    }
    const auto n1 = static_cast<std::size_t>(s.reduced_cols), m1 = static_cast<std::size_t>(s.reduced_rows);
    const std::vector<double> xr = tools::read_binary_vector<double>(dir / "x_f64.bin", n1);
    const std::vector<double> yr = tools::read_binary_vector<double>(dir / "y_f64.bin", m1);
    const std::vector<double> zr = tools::read_binary_vector<double>(dir / "z_f64.bin", n1);
    require_finite(xr, "reduced primal solution");
    require_finite(yr, "reduced row dual solution");
    require_finite(zr, "reduced bound dual solution");
    double* dx = tools::copy_to_device(xr);
    double* dy = tools::copy_to_device(yr);
    double* dz = tools::copy_to_device(zr);
    ps::GpuPostsolveResult post = ps::postsolve_gpu(dx, dy, dz, s.record, &dev.lp);
    const auto n0 = static_cast<std::size_t>(s.original_cols), m0 = static_cast<std::size_t>(s.original_rows);
    const std::vector<double> x = tools::copy_to_host(post.x_org, n0);
    const std::vector<double> y = tools::copy_to_host(post.y_org, m0);
    const std::vector<double> z = tools::copy_to_host(post.z_org, n0);
    for (double* p : {dx, dy, dz, post.x_org, post.y_org, post.z_org})
      if (p) (void)hipFree(p);
    const tools::PrimalFeasibility f = tools::compute_primal_feasibility(model.lp, x);
    std::cout << "status ok\n"
              << "reduced_rows " << s.reduced_rows << "\n"
              << "reduced_cols " << s.reduced_cols << "\n"
              << "primal_feas_l2_normalized " << f.normalized << "\n"
              << "row_violation_l2 " << f.row_l2 << "\n"
              << "row_violation_max " << f.row_max << "\n"
              << "col_violation_l2 " << f.col_l2 << "\n"
              << "col_violation_max " << f.col_max << "\n"
              << "record_hash " << tools::record_hash(s.record) << "\n"
              << "x_org_hash " << tools::raw_vector_hash(x) << "\n"
              << "y_org_hash " << tools::raw_vector_hash(y) << "\n"
              << "z_org_hash " << tools::raw_vector_hash(z) << "\n";
    if (std::getenv("GPUPRESOLVER_PRINT_TOP_VIOLATIONS") != nullptr) print_top_violations(model.lp, x, 10);
    return std::isfinite(f.normalized) && f.normalized <= 1.0e-6 ? 0 : 1;
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 2;
  }
}
