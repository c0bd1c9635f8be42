// Presolve, solve and postsolve of one LP in one call, with the reduced LP
// kept in device memory throughout.  The postsolve's preparation (tape
// upload and validation, the dependency-level replay schedules built on the
// host, the recovery steps) needs only the presolve record, not the
// solution, so it runs on a second thread and stream while the solver runs;
// the prepared replay then postsolves the solution.  Both halves are the
// library's ordinary entry points (run_gpu_presolve_with_reduced_lp,
// pdhg_solve, PostsolvePlan), so the result is bit for bit what the same
// steps give one after another.
#include <chrono>
#include <exception>
#include <memory>
#include <thread>

#include "gp_host.hpp"
#include "gpu_presolver/presolve/gpu_presolve.hpp"

namespace gpu_presolver::amd {

using presolve::GpuPresolveSummary;
using presolve::LPInfoGpu;
using presolve::PresolveParams;

namespace {

double since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

// Joins the preparing thread on every way out of the caller's scope.
struct Joiner {
  std::thread& t;
  ~Joiner() {
    if (t.joinable()) t.join();
  }
};

}  // namespace

PipelineReport presolve_solve(const LPInfoGpu& lp, const PresolveParams& prm, const PdhgOptions& so, double* x,
                              double* y, double* z) {
  AbandonOnThrow abandon_guard;
  const auto t0 = std::chrono::steady_clock::now();
  PipelineReport rep;
  GpuPresolveSummary sum = presolve::run_gpu_presolve_with_reduced_lp(lp, prm);
  struct FreeReduced {
    GpuPresolveSummary& s;
    ~FreeReduced() { presolve::free_gpu_presolve_reduced_lp(s); }
  } free_reduced{sum};
  rep.presolve_seconds = since(t0);
  const int32_t m0 = lp.A.rows, n0 = lp.A.cols;
  auto& c = ctx();
  if (sum.has_infeasible || sum.has_unbounded) {
    // Nothing to solve: the outputs are cleared and the flags reported.
    CopyBatch().zero(x, n0).zero(z, n0).zero(y, m0).run("pipeline outputs");
    c.sync("pipeline");
    rep.solve.status = -1;
    rep.total_seconds = since(t0);
    rep.copy_summary(sum);
    return rep;
  }
  // The second stream's context outlives the plan built on it.
  Ctx side;
  std::unique_ptr<PostsolvePlan> plan;
  std::exception_ptr side_error;
  double prepare_s = 0.0;
  std::thread prep([&] {
    try {
      ScopedCtx scope(side);
      const auto t = std::chrono::steady_clock::now();
      plan = std::make_unique<PostsolvePlan>(sum.record, &lp);
      side.sync("postsolve prepare");
      prepare_s = since(t);
    } catch (...) {
      side_error = std::current_exception();
      side.abandon();
    }
  });
  Joiner joiner{prep};
  const LPInfoGpu& red = sum.reduced_lp;
  DBuf<double> xr(red.A.cols, "reduced x"), yr(red.A.rows, "reduced y"), zr(red.A.cols, "reduced z");
  const auto t1 = std::chrono::steady_clock::now();
  rep.solve = pdhg_solve(red, so, xr, yr, zr);
  c.sync("pipeline solve");
  rep.solve_seconds = since(t1);
  prep.join();
  if (side_error) std::rethrow_exception(side_error);
  rep.prepare_seconds = prepare_s;
  const auto t2 = std::chrono::steady_clock::now();
  plan->enqueue(xr, yr, zr, x, y, z);
  c.sync("pipeline postsolve");
  rep.postsolve_seconds = since(t2);
  plan.reset();  // its buffers go back to the side stream, which is idle
  rep.copy_summary(sum);
  rep.total_seconds = since(t0);
  return rep;
}

void PipelineReport::copy_summary(const GpuPresolveSummary& s) {
  reduced_rows = s.reduced_rows;
  reduced_cols = s.reduced_cols;
  reduced_nnz = s.reduced_nnz;
  iterations = s.iterations;
  obj_constant_delta = s.obj_constant_delta;
  has_infeasible = s.has_infeasible;
  has_unbounded = s.has_unbounded;
}

}  // namespace gpu_presolver::amd
