// gpu_presolver_export_reduced <model.mps|-> <out_dir>: presolve on the GPU
// and write the reduced LP as an export directory (meta.txt and raw binary
// vectors, format gpu_presolver_reduced_lp_v1; reference
// cpp/tools/gpu_presolver_export_reduced.cpp).  Exit status 0, 1 when
// presolve reports the LP infeasible or unbounded, 2 on an error (including
// a failed write).
#include <filesystem>
#include <iostream>
#include <string>

#include "gpu_presolver/presolve/gpu_presolve.hpp"
#include "gpu_presolver/tools.hpp"

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: gpu_presolver_export_reduced <model.mps|-> <out_dir>\n";
    return 2;
  }
  namespace tools = gpu_presolver::tools;
  namespace ps = gpu_presolver::presolve;
  try {
    const std::filesystem::path out_dir = argv[2];
    mps_reader::MpsModel model = tools::read_mps_file(argv[1]);
    tools::DeviceLpOwner dev = tools::upload_lp(model.lp);
    ps::PresolveParams params;
    tools::apply_presolve_env_overrides(params);
    ps::GpuPresolveSummary s = ps::run_gpu_presolve_with_reduced_lp(dev.lp, params);
    try {
      tools::write_lp_export(out_dir, model.name, s);
    } catch (...) {
      ps::free_gpu_presolve_reduced_lp(s);
      throw;
    }
    std::cout << "status " << (s.has_infeasible ? "infeasible" : (s.has_unbounded ? "unbounded" : "ok")) << "\n"
              << "reduced_rows " << s.reduced_rows << "\n"
              << "reduced_cols " << s.reduced_cols << "\n"
              << "reduced_nnz " << s.reduced_nnz << "\n"
              << "out_dir " << out_dir.string() << "\n";
    const int rc = s.has_infeasible || s.has_unbounded ? 1 : 0;
    ps::free_gpu_presolve_reduced_lp(s);
    return rc;
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 2;
  }
}
