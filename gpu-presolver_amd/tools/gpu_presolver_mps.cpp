// gpu_presolver_mps <model.mps|->: read an MPS file, presolve it on the GPU
// and print the summary (reference cpp/tools/gpu_presolver_mps.cpp).  Exit
// status 0, 1 when presolve reports the LP infeasible or unbounded, 2 on an
// error.  GPUPRESOLVER_* environment variables override PresolveParams.
#include <chrono>
#include <iostream>
#include <string>

#include "gpu_presolver/presolve/gpu_presolve.hpp"
#include "gpu_presolver/tools.hpp"

int main(int argc, char** argv) {
  if (argc != 2) {
    std::cerr << "usage: gpu_presolver_mps <model.mps|->\n";
    return 2;
  }
  namespace tools = gpu_presolver::tools;
  namespace ps = gpu_presolver::presolve;
  using clock = std::chrono::steady_clock;
  try {
    const auto t0 = clock::now();
    mps_reader::MpsModel model = tools::read_mps_file(argv[1]);
    const auto t1 = clock::now();
    tools::DeviceLpOwner dev = tools::upload_lp(model.lp);
    ps::PresolveParams params;
    tools::apply_presolve_env_overrides(params);
    const auto t2 = clock::now();
    const ps::GpuPresolveSummary s = ps::run_gpu_presolve_with_record(dev.lp, params);
    const auto t3 = clock::now();
    const bool changed = s.reduced_rows != s.original_rows || s.reduced_cols != s.original_cols;
    std::cout << "model " << model.name << '\n'
              << "status " << (s.has_infeasible ? "infeasible" : (s.has_unbounded ? "unbounded" : "ok")) << '\n'
              << "changed " << (changed ? 1 : 0) << '\n'
              << "original_rows " << s.original_rows << '\n'
              << "original_cols " << s.original_cols << '\n'
              << "original_nnz " << dev.lp.A.nnz << '\n'
              << "reduced_rows " << s.reduced_rows << '\n'
              << "reduced_cols " << s.reduced_cols << '\n'
              << "reduced_nnz " << s.reduced_nnz << '\n'
              << "objective_shift " << s.obj_constant_delta << '\n'
              << "iterations " << s.iterations << '\n'
              << "parse_seconds " << std::chrono::duration<double>(t1 - t0).count() << '\n'
              << "presolve_seconds " << std::chrono::duration<double>(t3 - t2).count() << '\n';
    return s.has_infeasible || s.has_unbounded ? 1 : 0;
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 2;
  }
}
