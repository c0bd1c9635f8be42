// Postsolve of records that carry the structure-specific reductions' primal
// recovery steps, through the C++ API (postsolve_gpu,
// postsolve_restore_structural_primal_gpu).  The first four cases are the
// reference's own (cpp/tests/test_gpu_postsolve.cpp:868-1058); the others
// check each recovery kind with values worked out by hand.  Prints
// "ok <case>"; exits 1 at the first failure.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpu_presolver/presolve/gpu_postsolve.hpp"

namespace ps = gpu_presolver::presolve;
using ps::PostsolveDualMode;
using ps::PostsolveReductionType;
using ps::PrimalRecoveryCheckpoint;
using ps::PrimalRecoveryKind;
using ps::PresolveRecordGpu;

namespace {

const double kInf = std::numeric_limits<double>::infinity();

void require(bool ok, const std::string& what) {
  if (!ok) throw std::runtime_error(what);
}

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
T* dev(const std::vector<T>& v) {
  if (v.empty()) return nullptr;
  T* p = nullptr;
  hip_ok(hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * v.size()), "hipMalloc");
  hip_ok(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), "hipMemcpy");
  return p;
}

template <class T>
std::vector<T> host(const T* p, size_t n) {
  std::vector<T> v(n);
  if (n) hip_ok(hipMemcpy(v.data(), p, sizeof(T) * n, hipMemcpyDeviceToHost), "hipMemcpy");
  return v;
}

struct Result {
  std::vector<double> x, y, z;
};

Result run(const PresolveRecordGpu& rec, const std::vector<double>& xr, const std::vector<double>& zr,
           const ps::LPInfoGpu* orig = nullptr, const std::vector<double>& yr = {}, const double* original_l = nullptr) {
  double* dx = dev(xr);
  double* dy = dev(yr);
  double* dz = dev(zr);
  ps::GpuPostsolveResult r = ps::postsolve_gpu(dx, dy, dz, rec, orig, original_l);
  Result out{host(r.x_org, static_cast<size_t>(rec.n0)), host(r.y_org, static_cast<size_t>(rec.m0)),
             host(r.z_org, static_cast<size_t>(rec.n0))};
  for (void* p : {static_cast<void*>(dx), static_cast<void*>(dy), static_cast<void*>(dz), static_cast<void*>(r.x_org),
                  static_cast<void*>(r.y_org), static_cast<void*>(r.z_org)})
    if (p) (void)hipFree(p);
  return out;
}

void near(double a, double b, const char* what) {
  require(std::fabs(a - b) < 1e-12, std::string(what) + ": " + std::to_string(a) + " != " + std::to_string(b));
}

ps::StructuralL1PrimalRecoveryStep split_step(int t, int e, double rho) {
  ps::StructuralL1PrimalRecoveryStep s;
  s.pattern = "split";
  s.splits.push_back(ps::StructuralL1SplitRecovery{t, e, rho});
  return s;
}

// Reference test_dual_tape_replay_runs_after_structural_primal_recovery.
void dual_replay_after_structural_recovery() {
  ps::LPInfoGpu o;
  o.A = {1, 2, 2, dev<int32_t>({0, 2}), dev<int32_t>({0, 1}), dev<double>({1.0, 1.0})};
  o.AT = {2, 1, 2, dev<int32_t>({0, 1, 2}), dev<int32_t>({0, 0}), dev<double>({1.0, 1.0})};
  o.c = dev<double>({5.0, 0.0});
  o.l = dev<double>({3.5, -kInf});
  o.u = dev<double>({kInf, kInf});
  o.AL = dev<double>({-kInf});
  o.AU = dev<double>({kInf});
  PresolveRecordGpu rec;
  rec.m0 = rec.m1 = 1;
  rec.n0 = rec.n1 = 2;
  rec.row_red2org = {0};
  rec.col_red2org = {0, 1};
  rec.tape.types = {static_cast<int32_t>(PostsolveReductionType::BoundChangeNoRow)};
  rec.tape.index_starts = {0, 1};
  rec.tape.value_starts = {0, 0};
  rec.tape.dual_modes = {static_cast<uint8_t>(PostsolveDualMode::Minimal)};
  rec.tape.indices = {0};
  rec.structural_primal_recoveries.push_back(split_step(0, 1, 1.0));
  rec.primal_recovery_timeline.push_back(PrimalRecoveryCheckpoint{PrimalRecoveryKind::StructuralL1, 0, 0});
  const Result r = run(rec, {3.0, 1.0}, {99.0, 0.0}, &o, {0.0});
  near(r.x[0], 4.0, "x0");
  near(r.x[1], 2.0, "x1");
  near(r.z[0], 0.0, "z0");
  for (void* p : {static_cast<void*>(o.A.rowPtr), static_cast<void*>(o.A.colVal), static_cast<void*>(o.A.nzVal),
                  static_cast<void*>(o.AT.rowPtr), static_cast<void*>(o.AT.colVal), static_cast<void*>(o.AT.nzVal),
                  static_cast<void*>(o.c), static_cast<void*>(o.l), static_cast<void*>(o.u), static_cast<void*>(o.AL),
                  static_cast<void*>(o.AU)})
    (void)hipFree(p);
}

PresolveRecordGpu fme_record(int n0, std::vector<int32_t> red2org) {
  PresolveRecordGpu rec;
  rec.n0 = n0;
  rec.n1 = static_cast<int32_t>(red2org.size());
  rec.col_red2org = std::move(red2org);
  rec.tape.types = {static_cast<int32_t>(PostsolveReductionType::FmeCol)};
  rec.tape.index_starts = {0, 4};
  rec.tape.value_starts = {0, 5};
  rec.tape.dual_modes = {static_cast<uint8_t>(PostsolveDualMode::Minimal)};
  rec.tape.indices = {0, 2, 1, 0};  // no deleted rows, eliminate col 2, one support column: 0
  rec.tape.vals = {1.0, 1.0, 1.0, 0.0, 1.0};  // one lower affine side: x2 >= x0
  return rec;
}

// Reference test_tape_checkpoint_replays_after_structural_primal.
void tape_checkpoint_after_structural_primal() {
  PresolveRecordGpu rec = fme_record(3, {0, 1});
  rec.structural_primal_recoveries.push_back(split_step(0, 1, 1.0));
  rec.primal_recovery_timeline.push_back(PrimalRecoveryCheckpoint{PrimalRecoveryKind::StructuralL1, 0, 1});
  const Result r = run(rec, {3.0, 1.0}, {0.0, 0.0});
  near(r.x[0], 4.0, "x0");
  near(r.x[1], 2.0, "x1");
  near(r.x[2], 4.0, "x2");
}

// Reference test_structural_tape_structural_timeline_replays_exact_reverse_order.
void timeline_in_exact_reverse_order() {
  PresolveRecordGpu rec = fme_record(4, {0, 1, 3});
  rec.structural_primal_recoveries.push_back(split_step(0, 1, 1.0));
  rec.primal_recovery_timeline.push_back(PrimalRecoveryCheckpoint{PrimalRecoveryKind::StructuralL1, 0, 0});
  rec.structural_primal_recoveries.push_back(split_step(0, 3, 1.0));
  rec.primal_recovery_timeline.push_back(PrimalRecoveryCheckpoint{PrimalRecoveryKind::StructuralL1, 1, 1});
  const Result r = run(rec, {2.0, 1.0, 1.0}, {0.0, 0.0, 0.0});
  require(r.x == std::vector<double>({4.0, 2.0, 3.0, 1.0}), "x");
}

template <class F>
bool throws(F f) {
  try {
    f();
  } catch (const std::runtime_error&) {
    return true;
  }
  return false;
}

// Reference test_postsolve_rejects_invalid_primal_recovery_checkpoint, plus
// the other malformed timelines and payloads.
void rejects_invalid_timelines() {
  PresolveRecordGpu rec;
  rec.n0 = 2;
  rec.structural_primal_recoveries.push_back(split_step(0, 1, 1.0));
  rec.primal_recovery_timeline.push_back(PrimalRecoveryCheckpoint{PrimalRecoveryKind::StructuralL1, 0, 1});
  require(throws([&] { ps::postsolve_gpu(nullptr, nullptr, nullptr, rec); }), "checkpoint past the tape");
  rec.primal_recovery_timeline[0].tape_position = 0;
  rec.primal_recovery_timeline.push_back(PrimalRecoveryCheckpoint{PrimalRecoveryKind::StructuralL1, 0, 0});
  require(throws([&] { ps::postsolve_gpu(nullptr, nullptr, nullptr, rec); }), "duplicate checkpoint");
  rec.primal_recovery_timeline.pop_back();
  rec.structural_primal_recoveries.push_back(split_step(0, 1, 1.0));
  require(throws([&] { ps::postsolve_gpu(nullptr, nullptr, nullptr, rec); }), "payload without a checkpoint");
  rec.structural_primal_recoveries.pop_back();
  rec.structural_primal_recoveries[0].splits[0].e_col = 2;
  require(throws([&] { ps::postsolve_gpu(nullptr, nullptr, nullptr, rec); }), "column out of range");
  PresolveRecordGpu a;
  a.n0 = 2;
  ps::AntipodalComponentPrimalRecoveryStep st;
  st.plus_cols = {0};
  st.minus_cols = {1};
  a.antipodal_component_recoveries.push_back(st);
  a.primal_recovery_timeline.push_back(PrimalRecoveryCheckpoint{PrimalRecoveryKind::AntipodalComponent, 0, 0});
  require(throws([&] { ps::postsolve_gpu(nullptr, nullptr, nullptr, a); }), "antipodal payload sizes");
}

// Each structural kind in its order: splits, outer pairs, linked slacks,
// max slacks.  x = (1.5, 2, -1, 4, 0.5, 3); split (0, 1, rho 2) gives
// x0 = 3.5, x1 = -0.25; outer (3 -> 2) x2 = 4; linked (4 = 0.5 x0) x4 = 1.75;
// max slack 5 over (2 x1, -1 x3) = max(lower, -0.5, -4).
void structural_kinds_in_order() {
  PresolveRecordGpu rec;
  rec.n0 = rec.n1 = 6;
  rec.col_red2org = {0, 1, 2, 3, 4, 5};
  ps::StructuralL1PrimalRecoveryStep st;
  st.pattern = "kinds";
  st.splits.push_back(ps::StructuralL1SplitRecovery{0, 1, 2.0});
  st.outer_pairs.push_back(ps::StructuralOuterPairRecovery{3, 2});
  st.linked_slacks.push_back(ps::StructuralLinkedSlackRecovery{4, 0, 0.5});
  ps::StructuralMaxSlackRecovery mx;
  mx.slack_col = 5;
  mx.t_cols = {1, 3};
  mx.factors = {2.0, -1.0};
  st.max_slacks.push_back(mx);
  rec.structural_primal_recoveries.push_back(st);
  rec.primal_recovery_timeline.push_back(PrimalRecoveryCheckpoint{PrimalRecoveryKind::StructuralL1, 0, 0});
  const std::vector<double> xr = {1.5, 2.0, -1.0, 4.0, 0.5, 3.0}, zr(6, 0.0);
  const Result r0 = run(rec, xr, zr);
  require(r0.x == std::vector<double>({3.5, -0.25, 4.0, 4.0, 1.75, 0.0}), "no lower bounds");
  double* l = dev<double>({0.0, 0.0, 0.0, 0.0, 0.0, -10.0});
  const Result r1 = run(rec, xr, zr, nullptr, {}, l);
  require(r1.x == std::vector<double>({3.5, -0.25, 4.0, 4.0, 1.75, -0.5}), "finite lower bound");
  (void)hipFree(l);
  // the stand-alone entry point applies the same step in place
  double* x = dev<double>(xr);
  ps::postsolve_restore_structural_primal_gpu(x, rec.structural_primal_recoveries, nullptr);
  require(host(x, 6) == std::vector<double>({3.5, -0.25, 4.0, 4.0, 1.75, 0.0}), "stand-alone recovery");
  (void)hipFree(x);
}

// Antipodal members from their roots, the roots read before any member is
// written: x = (2, 7, 0, 0, 1); plus (0, 2) from roots (2, 0) with lower 1,
// minus (3, 4) from root 1 with lower (0, 0.5).
void antipodal_component() {
  PresolveRecordGpu rec;
  rec.n0 = rec.n1 = 5;
  rec.col_red2org = {0, 1, 2, 3, 4};
  ps::AntipodalComponentPrimalRecoveryStep st;
  st.plus_cols = {0, 2};
  st.minus_cols = {3, 4};
  st.root_plus_cols = {2, 0};
  st.root_minus_cols = {1, 1};
  st.plus_lower = {1.0, 1.0};
  st.minus_lower = {0.0, 0.5};
  rec.antipodal_component_recoveries.push_back(st);
  rec.primal_recovery_timeline.push_back(PrimalRecoveryCheckpoint{PrimalRecoveryKind::AntipodalComponent, 0, 0});
  const Result r = run(rec, {2.0, 7.0, 0.0, 0.0, 1.0}, std::vector<double>(5, 0.0));
  require(r.x == std::vector<double>({1.0, 7.0, 3.0, 7.0, 7.5}), "antipodal members");
}

}  // namespace

int main() {
  const std::vector<std::pair<const char*, std::function<void()>>> cases = {
      {"dual_replay_after_structural_recovery", dual_replay_after_structural_recovery},
      {"tape_checkpoint_after_structural_primal", tape_checkpoint_after_structural_primal},
      {"timeline_in_exact_reverse_order", timeline_in_exact_reverse_order},
      {"rejects_invalid_timelines", rejects_invalid_timelines},
      {"structural_kinds_in_order", structural_kinds_in_order},
      {"antipodal_component", antipodal_component},
  };
  for (const auto& [name, fn] : cases) {
    try {
      fn();
    } catch (const std::exception& e) {
      std::printf("FAIL %s: %s\n", name, e.what());
      return 1;
    }
    std::printf("ok %s\n", name);
  }
  return 0;
}
