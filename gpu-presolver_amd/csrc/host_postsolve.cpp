// Postsolve driver (reference cpp/src/presolve/gpu_postsolve.cu:1620-1721):
// scatter of the reduced solution, fixed-value restore, reverse tape replay
// (primal pass, then dual pass) and the final column-dual projection.
#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <vector>

#include "gp_host.hpp"
#include "gp_record_kernels.h"
#include "gpu_presolver/presolve/gpu_postsolve.hpp"

namespace gpu_presolver::presolve {

using amd::check;
using amd::ctx;
using amd::DBuf;

namespace {

constexpr double kPostsolveTol = 1.0e-7;

double* malloc_f64(int32_t n, const char* what) {
  if (n <= 0) return nullptr;
  void* p = nullptr;
  check(hipMalloc(&p, sizeof(double) * static_cast<size_t>(n)), what);
  return static_cast<double*>(p);
}

// Tape upload (upload_postsolve_tape, gpu_postsolve.cu:1060-1120).
struct DeviceTape {
  DBuf<int32_t> types, index_starts, value_starts, indices;
  DBuf<double> vals;
  gp_tape_dev view{};
  explicit DeviceTape(const PostsolveTape& t) {
    auto& c = ctx();
    const int32_t count = static_cast<int32_t>(t.types.size());
    if (t.index_starts.size() != t.types.size() + 1 || t.value_starts.size() != t.types.size() + 1 ||
        (!t.types.empty() && (t.index_starts.back() != static_cast<int32_t>(t.indices.size()) ||
                              t.value_starts.back() != static_cast<int32_t>(t.vals.size()))))
      throw std::runtime_error("postsolve tape offsets are inconsistent");
    types.alloc(count);
    index_starts.alloc(static_cast<int64_t>(count) + 1);
    value_starts.alloc(static_cast<int64_t>(count) + 1);
    indices.alloc(static_cast<int64_t>(t.indices.size()));
    vals.alloc(static_cast<int64_t>(t.vals.size()));
    c.to_device(types.get(), t.types.data(), count, "tape types upload");
    c.to_device(index_starts.get(), t.index_starts.data(), count + 1, "tape index starts upload");
    c.to_device(value_starts.get(), t.value_starts.data(), count + 1, "tape value starts upload");
    c.to_device(indices.get(), t.indices.data(), static_cast<int64_t>(t.indices.size()), "tape indices upload");
    c.to_device(vals.get(), t.vals.data(), static_cast<int64_t>(t.vals.size()), "tape values upload");
    view = gp_tape_dev{count, types, index_starts, value_starts, indices, vals};
  }
};

gp_orig_dev orig_view(const LPInfoGpu* o) {
  gp_orig_dev v{};
  if (o == nullptr) return v;
  v.present = 1;
  v.a_ptr = o->A.rowPtr;
  v.a_idx = o->A.colVal;
  v.a_val = o->A.nzVal;
  v.at_ptr = o->AT.rowPtr;
  v.at_idx = o->AT.colVal;
  v.at_val = o->AT.nzVal;
  v.c = o->c;
  v.l = o->l;
  v.u = o->u;
  return v;
}

void upload_map(DBuf<int32_t>& dst, const std::vector<int32_t>& src, int32_t bound, const char* what) {
  for (int32_t v : src)
    if (v < 0 || v >= bound) throw std::runtime_error(std::string(what) + ": index out of range");
  dst.alloc(static_cast<int64_t>(src.size()));
  ctx().to_device(dst.get(), src.data(), static_cast<int64_t>(src.size()), what);
}

}  // namespace

void postsolve_restore_structural_primal_gpu(double* x_org, const std::vector<StructuralL1PrimalRecoveryStep>& recoveries,
                                             const double* original_l) {
  (void)x_org;
  (void)original_l;
  if (!recoveries.empty())
    throw std::runtime_error(
        "postsolve_restore_structural_primal_gpu: structural L1 recoveries are produced by the structure-specific "
        "rules, which this library does not implement");
  check(hipDeviceSynchronize(), "postsolve_restore_structural_primal_gpu synchronize");
}

GpuPostsolveResult postsolve_gpu(double* x_red, double* y_red, double* z_red, const PresolveRecordGpu& record,
                                 const double* original_l) {
  return postsolve_gpu(x_red, y_red, z_red, record, nullptr, original_l);
}

GpuPostsolveResult postsolve_gpu(double* x_red, double* y_red, double* z_red, const PresolveRecordGpu& record,
                                 const LPInfoGpu* original_model, const double* original_l) {
  (void)original_l;
  if (!record.primal_recovery_timeline.empty() || !record.structural_primal_recoveries.empty() ||
      !record.antipodal_component_recoveries.empty())
    throw std::runtime_error(
        "postsolve_gpu: the record holds structure-specific recovery steps, which this library does not implement");
  if (static_cast<int32_t>(record.col_red2org.size()) != record.n1 ||
      static_cast<int32_t>(record.row_red2org.size()) != record.m1)
    throw std::runtime_error("postsolve_gpu: record maps do not match m1/n1");
  if (record.fixed_idx.size() != record.fixed_val.size())
    throw std::runtime_error("postsolve_gpu: fixed_idx and fixed_val differ in length");
  if (original_model != nullptr &&
      (original_model->A.rows != record.m0 || original_model->A.cols != record.n0))
    throw std::runtime_error("postsolve_gpu: original model does not match the record's dimensions");
  check(hipDeviceSynchronize(), "postsolve entry synchronize");
  auto& c = ctx();
  hipStream_t s = c.stream();
  GpuPostsolveResult res;
  res.n0 = record.n0;
  res.m0 = record.m0;
  const int32_t n0 = record.n0, m0 = record.m0;
  try {
    res.x_org = malloc_f64(n0, "hipMalloc postsolve x_org");
    res.z_org = malloc_f64(n0, "hipMalloc postsolve z_org");
    res.y_org = malloc_f64(m0, "hipMalloc postsolve y_org");
    DBuf<uint8_t> zret(n0);
    DBuf<int32_t> col_map, row_map, fidx;
    DBuf<double> fval;
    upload_map(col_map, record.col_red2org, n0, "col_red2org upload");
    upload_map(row_map, record.row_red2org, m0, "row_red2org upload");
    upload_map(fidx, record.fixed_idx, n0, "fixed_idx upload");
    const int32_t nf = static_cast<int32_t>(record.fixed_val.size());
    fval.alloc(nf);
    c.to_device(fval.get(), record.fixed_val.data(), nf, "fixed_val upload");
    if (n0 > 0) {
      check(hipMemsetAsync(res.x_org, 0, sizeof(double) * n0, s), "x_org clear");
      check(hipMemsetAsync(res.z_org, 0, sizeof(double) * n0, s), "z_org clear");
      check(hipMemsetAsync(zret.get(), 0, n0, s), "z_retrieved clear");
    }
    if (m0 > 0) check(hipMemsetAsync(res.y_org, 0, sizeof(double) * m0, s), "y_org clear");
    if (x_red) check(gpk_scatter(x_red, col_map, record.n1, res.x_org, s), "scatter x");
    if (y_red && m0 > 0) check(gpk_scatter(y_red, row_map, record.m1, res.y_org, s), "scatter y");
    if (z_red) check(gpk_scatter(z_red, col_map, record.n1, res.z_org, s), "scatter z");
    check(gpk_mark(col_map, record.n1, zret, s), "mark retrieved");
    check(gpk_restore_fixed(fidx, fval, nf, res.x_org, s), "restore fixed");
    DeviceTape tape(record.tape);
    const gp_orig_dev orig = orig_view(original_model);
    double* y = m0 > 0 ? res.y_org : nullptr;
    check(gpk_replay(tape.view, orig, res.x_org, y, res.z_org, zret, 0, tape.view.count, 1, 0, kPostsolveTol, s),
          "replay primal");
    check(gpk_replay(tape.view, orig, res.x_org, y, res.z_org, zret, 0, tape.view.count, 0, 1, kPostsolveTol, s),
          "replay dual");
    // project_column_duals_from_original (gpu_postsolve.cu:1370-1435)
    if (original_model != nullptr && y != nullptr && original_model->AT.rows > 0 && original_model->c &&
        original_model->l && original_model->u) {
      DBuf<uint8_t> prot(n0);
      check(hipMemsetAsync(prot.get(), 0, n0, s), "protected columns clear");
      check(gpk_protect_columns(tape.view, n0, prot, s), "protected columns");
      check(gpk_project_duals(orig, prot, res.x_org, y, res.z_org, n0, kPostsolveTol, s), "project duals");
    }
    c.sync("postsolve_gpu synchronize");
  } catch (...) {
    (void)hipStreamSynchronize(s);
    if (res.x_org) (void)hipFree(res.x_org);
    if (res.y_org) (void)hipFree(res.y_org);
    if (res.z_org) (void)hipFree(res.z_org);
    throw;
  }
  return res;
}

}  // namespace gpu_presolver::presolve
