// Presolve driver (reference cpp/src/presolve/gpu_presolve.cu): the working
// LP in HBM, per-phase plans and statistics, plan application (compaction,
// re-transpose, record update) and the tiered and fixed schedulers.
//
// Device-side layout: the working LP owns one CSR pair (A, AT) and five
// vectors from the stream-ordered pool; each phase owns one contiguous plan
// block (masks + five vectors) and one contiguous statistics block.  The
// reduced-to-original maps of the record stay on the device for the whole
// run and are copied to the host once at the end.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <optional>
#include <type_traits>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "gp_host.hpp"
#include "gp_record_kernels.h"
#include "gpu_presolver/presolve/gpu_presolve.hpp"
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

namespace gpu_presolver::amd {
namespace {
// kMinNnzForBatchedClosure (gpu_presolve.cu:2506).
std::atomic<int32_t> g_closure_min_nnz{10000000};
thread_local int32_t t_fail_after_updates = 0;
thread_local bool t_check_plan_reuse = false;
thread_local int64_t t_plan_reuses = 0;
}  // namespace
void count_plan_reuse() { ++t_plan_reuses; }
int64_t plan_reuses() { return t_plan_reuses; }
void set_batched_closure_min_nnz(int32_t nnz) { g_closure_min_nnz.store(nnz); }
void check_plan_reuse(bool on) { t_check_plan_reuse = on; }
bool plan_reuse_checked() { return t_check_plan_reuse; }
void fail_after_record_updates(int32_t count) { t_fail_after_updates = count > 0 ? count : 0; }
// Throws on the armed record update (after its host copies are queued).
void record_update_checkpoint() {
  if (t_fail_after_updates > 0 && --t_fail_after_updates == 0)
    throw std::runtime_error("presolve: injected failure after a record update");
}
int32_t batched_closure_min_nnz() { return g_closure_min_nnz.load(); }
}  // namespace gpu_presolver::amd

namespace gpu_presolver::presolve {

using amd::check;
using amd::ctx;
using amd::DBuf;
using RT = PostsolveReductionType;

namespace {

size_t align256(size_t b) { return (b + 255) & ~size_t{255}; }

// ---- working LP ------------------------------------------------------------
// Borrows the caller's LP until the first structural change, then owns
// pool-allocated arrays (gpu_presolve.cu:2021-2147, WorkingLp).
struct WorkLp {
  LPInfoGpu lp;
  bool owns = false;
  WorkLp() = default;
  WorkLp(const WorkLp&) = delete;
  WorkLp& operator=(const WorkLp&) = delete;
  ~WorkLp() { release(); }
  void release() noexcept {
    if (!owns) return;
    hipStream_t s = ctx().stream();
    amd::free_csr(lp.A);
    amd::free_csr(lp.AT);
    amd::dev_free(lp.c, s);
    amd::dev_free(lp.l, s);
    amd::dev_free(lp.u, s);
    amd::dev_free(lp.AL, s);
    amd::dev_free(lp.AU, s);
    lp = LPInfoGpu{};
    owns = false;
  }
};

double* pool_f64(int64_t n, const char* what) {
  return static_cast<double*>(amd::dev_alloc(sizeof(double) * static_cast<size_t>(n > 0 ? n : 1), ctx().stream(), what));
}

template <class T>
void d2d(T* dst, const T* src, int64_t n, const char* what) {
  if (n > 0)
    check(hipMemcpyAsync(dst, src, sizeof(T) * static_cast<size_t>(n), hipMemcpyDeviceToDevice, ctx().stream()), what);
}

// ---- phase plan ------------------------------------------------------------
// Fresh plan over the working LP: masks all ones, copies of c/l/u/AL/AU
// (gpu_presolve.cu:197-247), in one contiguous block.
struct PhasePlan {
  PresolvePlanGpu p;
  // init = false leaves the vectors to PhaseStats(lp, &plan), which fills
  // them in the same launch as the statistics.
  explicit PhasePlan(const LPInfoGpu& lp, bool init = true) {
    const size_t m = static_cast<size_t>(lp.A.rows), n = static_cast<size_t>(lp.A.cols);
    const size_t bytes = align256(8 * n) * 3 + align256(8 * m) * 2 + align256(m) + align256(n);
    char* b = static_cast<char*>(amd::dev_alloc(bytes, ctx().stream(), "plan storage"));
    p.contiguous_storage = b;
    p.new_c = reinterpret_cast<double*>(b);
    b += align256(8 * n);
    p.new_l = reinterpret_cast<double*>(b);
    b += align256(8 * n);
    p.new_u = reinterpret_cast<double*>(b);
    b += align256(8 * n);
    p.new_AL = reinterpret_cast<double*>(b);
    b += align256(8 * m);
    p.new_AU = reinterpret_cast<double*>(b);
    b += align256(8 * m);
    p.keep_row_mask = reinterpret_cast<uint8_t*>(b);
    b += align256(m);
    p.keep_col_mask = reinterpret_cast<uint8_t*>(b);
    if (init)
      check(gpk_plan_init(lp.A.rows, lp.A.cols, amd::plan_view(p), lp.c, lp.l, lp.u, lp.AL, lp.AU, ctx().stream()),
            "plan init");
  }
  ~PhasePlan() {
    if (p.has_new_A) amd::free_csr(p.new_A);
    amd::dev_free(p.contiguous_storage, ctx().stream());
  }
  PhasePlan(const PhasePlan&) = delete;
  PhasePlan& operator=(const PhasePlan&) = delete;
};

// ---- phase statistics ------------------------------------------------------
// recompute_row_stats_from_csr / recompute_col_stats_from_csr
// (gpu_presolve.cu:249-276, 696-732).
struct PhaseStats {
  PresolveStatsGpu s;
  // fresh_plan: a PhasePlan built with init = false, initialised here too.
  explicit PhaseStats(const LPInfoGpu& lp, PresolvePlanGpu* fresh_plan = nullptr) {
    const size_t m = static_cast<size_t>(lp.A.rows), n = static_cast<size_t>(lp.A.cols);
    const size_t bytes = align256(4 * m) * 2 + align256(8 * m) + align256(m) + align256(4 * n) * 2 +
                         align256(8 * n) + align256(n) * 2;
    char* b = static_cast<char*>(amd::dev_alloc(bytes, ctx().stream(), "stats storage"));
    s.contiguous_storage = b;
    s.row_nnz = reinterpret_cast<int32_t*>(b);
    b += align256(4 * m);
    s.singleton_row_col = reinterpret_cast<int32_t*>(b);
    b += align256(4 * m);
    s.singleton_row_val = reinterpret_cast<double*>(b);
    b += align256(8 * m);
    s.singleton_row_mask = reinterpret_cast<uint8_t*>(b);
    b += align256(m);
    s.col_nnz = reinterpret_cast<int32_t*>(b);
    b += align256(4 * n);
    s.column_singleton_row = reinterpret_cast<int32_t*>(b);
    b += align256(4 * n);
    s.column_singleton_val = reinterpret_cast<double*>(b);
    b += align256(8 * n);
    s.empty_col_mask = reinterpret_cast<uint8_t*>(b);
    b += align256(n);
    s.column_singleton_mask = reinterpret_cast<uint8_t*>(b);
    hipStream_t st = ctx().stream();
    if (fresh_plan) {
      const gpk_phase_stats v{s.row_nnz, s.singleton_row_mask, s.singleton_row_col, s.singleton_row_val, s.col_nnz,
                              s.empty_col_mask, s.column_singleton_mask, s.column_singleton_row,
                              s.column_singleton_val};
      check(gpk_phase_init(lp.A.rows, lp.A.cols, amd::plan_view(*fresh_plan), lp.c, lp.l, lp.u, lp.AL, lp.AU,
                           amd::csr_view(lp.A), amd::csr_view(lp.AT), v, st),
            "phase init");
      return;
    }
    if (m > 0)
      check(gpk_csr_stats(lp.A.rowPtr, lp.A.colVal, lp.A.nzVal, lp.A.rows, s.row_nnz, nullptr, s.singleton_row_mask,
                          s.singleton_row_col, s.singleton_row_val, st),
            "row stats");
    if (n > 0)
      check(gpk_csr_stats(lp.AT.rowPtr, lp.AT.colVal, lp.AT.nzVal, lp.AT.rows, s.col_nnz, s.empty_col_mask,
                          s.column_singleton_mask, s.column_singleton_row, s.column_singleton_val, st),
            "col stats");
  }
  ~PhaseStats() { amd::dev_free(s.contiguous_storage, ctx().stream()); }
  PhaseStats(const PhaseStats&) = delete;
  PhaseStats& operator=(const PhaseStats&) = delete;
};

// A phase that changes nothing leaves its plan as a fresh plan of the same
// LP (every flag clear, no tape, no objective delta): the next phase takes
// it and its statistics over instead of building them again.
bool plan_pristine(const PresolvePlanGpu& p) {
  return !p.has_change && !p.has_row_action && !p.has_col_action && !p.has_new_A && !p.has_infeasible &&
         !p.has_unbounded && p.obj_constant_delta == 0.0 && p.tape.types.empty() &&
         !p.has_structural_primal_recovery && !p.has_antipodal_component_recovery &&
         !p.has_covering_cost_dominance_reduction && !p.has_projected_auxiliary_reduction;
}

// Testing hook (gp_test_check_plan_reuse): a kept plan and its statistics
// against a freshly built pair, compared bit for bit where the rules read
// them (the singleton entries only where their masks are set).
void verify_kept_plan(const PresolvePlanGpu& kp, const PresolveStatsGpu& ks, const LPInfoGpu& lp) {
  PhasePlan fp(lp, false);
  PhaseStats fs(lp, &fp.p);
  auto& c = ctx();
  const int64_t m = lp.A.rows, n = lp.A.cols;
  auto same = [&](const auto* a, const auto* b, int64_t count, const char* what, const uint8_t* only = nullptr) {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(a)>>;
    const std::vector<T> ha = c.to_host(a, count, what), hb = c.to_host(b, count, what);
    std::vector<uint8_t> mask;
    if (only) mask = c.to_host(only, count, what);
    for (int64_t i = 0; i < count; ++i) {
      if (only && !mask[static_cast<size_t>(i)]) continue;
      if (std::memcmp(&ha[static_cast<size_t>(i)], &hb[static_cast<size_t>(i)], sizeof(T)) != 0)
        throw std::logic_error(std::string("presolve: a kept phase plan differs from a fresh one in ") + what +
                               " at " + std::to_string(i));
    }
  };
  same(kp.new_c, fp.p.new_c, n, "c");
  same(kp.new_l, fp.p.new_l, n, "l");
  same(kp.new_u, fp.p.new_u, n, "u");
  same(kp.new_AL, fp.p.new_AL, m, "AL");
  same(kp.new_AU, fp.p.new_AU, m, "AU");
  same(kp.keep_row_mask, fp.p.keep_row_mask, m, "keep_row");
  same(kp.keep_col_mask, fp.p.keep_col_mask, n, "keep_col");
  same(ks.row_nnz, fs.s.row_nnz, m, "row_nnz");
  same(ks.singleton_row_mask, fs.s.singleton_row_mask, m, "singleton_row_mask");
  same(ks.singleton_row_col, fs.s.singleton_row_col, m, "singleton_row_col", fs.s.singleton_row_mask);
  same(ks.singleton_row_val, fs.s.singleton_row_val, m, "singleton_row_val", fs.s.singleton_row_mask);
  same(ks.col_nnz, fs.s.col_nnz, n, "col_nnz");
  same(ks.empty_col_mask, fs.s.empty_col_mask, n, "empty_col_mask");
  same(ks.column_singleton_mask, fs.s.column_singleton_mask, n, "column_singleton_mask");
  same(ks.column_singleton_row, fs.s.column_singleton_row, n, "column_singleton_row", fs.s.column_singleton_mask);
  same(ks.column_singleton_val, fs.s.column_singleton_val, n, "column_singleton_val", fs.s.column_singleton_mask);
}

// ---- record state ----------------------------------------------------------
struct RecordState {
  PresolveRecordGpu* rec = nullptr;
  DBuf<int32_t> row_r2o, col_r2o;  // device reduced-to-original maps
  int32_t m = 0, n = 0;            // their current lengths
};

// Index kinds of one phase tape for globalization: 1 row, 2 column, 0 none
// (append_globalized_postsolve_tape, gpu_presolve.cu:1363-1533).
void tape_index_kinds(const PostsolveTape& t, std::vector<uint8_t>& kind) {
  kind.assign(t.indices.size(), 0);
  const int32_t* I = t.indices.data();
  for (size_t k = 0; k < t.types.size(); ++k) {
    const int32_t b = t.index_starts[k];
    const int32_t cnt = t.index_starts[k + 1] - b;
    uint8_t* kd = kind.data() + b;
    const int32_t* ix = I + b;
    switch (static_cast<RT>(t.types[k])) {
      case RT::FixedCol:
        if (cnt > 0) {
          kd[0] = 2;
          for (int32_t q = 1; q < cnt; ++q) kd[q] = 1;
        }
        break;
      case RT::FixedColInf:
        if (cnt >= 2) {
          kd[1] = 2;
          int32_t q = 2;
          while (q < cnt) {
            const int32_t len = ix[q++];
            for (int32_t e = 0; e < len && q < cnt; ++e, ++q) kd[q] = 2;
          }
        }
        break;
      case RT::SubCol:
        if (cnt >= 3) {
          kd[0] = 2;
          kd[1] = 1;
          for (int32_t e = 0; e < ix[2] && 3 + e < cnt; ++e) kd[3 + e] = 2;
        }
        break;
      case RT::DuplicateColumn:
        if (cnt >= 2) kd[0] = kd[1] = 2;
        break;
      case RT::DuplicateRow:
        if (cnt >= 2) kd[0] = kd[1] = 1;
        break;
      case RT::DeletedRow:
        if (cnt >= 1) {
          kd[0] = 1;
          if (cnt >= 2) kd[1] = 2;
        }
        break;
      case RT::LhsChange:
      case RT::RhsChange:
      case RT::EqToIneq:
        if (cnt >= 1) kd[0] = 1;
        break;
      case RT::BoundChangeNoRow:
        if (cnt >= 1) kd[0] = 2;
        break;
      case RT::BoundChangeTheRow:
        if (cnt >= 2) {
          kd[0] = 2;
          kd[1] = 1;
        }
        break;
      case RT::DoubletonEquation:
        if (cnt >= 4) {
          kd[0] = kd[1] = 2;
          kd[2] = 1;
          for (int32_t e = 0; e < ix[3] && 4 + e < cnt; ++e) kd[4 + e] = 1;
        }
        break;
      case RT::FmeCol:
        if (cnt >= 3) {
          const int32_t inv = ix[0];
          kd[1] = 2;
          for (int32_t e = 0; e < inv && 2 + e < cnt; ++e) kd[2 + e] = 1;
          const int32_t sp = 2 + inv;
          if (sp < cnt)
            for (int32_t e = 0; e < ix[sp] && sp + 1 + e < cnt; ++e) kd[sp + 1 + e] = 2;
        }
        break;
      default:
        break;
    }
  }
}

// Tape chunks up to this size are globalized in the host-mapped staging area.
constexpr size_t kStagedTapeMaxBytes = size_t{32} << 20;

void append_global_tape(RecordState& rs, const PostsolveTape& src) {
  if (src.types.empty()) return;
  PostsolveTape& dst = rs.rec->tape;
  const int64_t cnt = static_cast<int64_t>(src.indices.size());
  const int32_t base_idx = static_cast<int32_t>(dst.indices.size());
  const int32_t base_val = static_cast<int32_t>(dst.vals.size());
  dst.indices.resize(dst.indices.size() + static_cast<size_t>(cnt));
  if (cnt > 0) {
    std::vector<uint8_t> kind;
    tape_index_kinds(src, kind);
    auto& c = ctx();
    // The indices and their kinds go to host-mapped staging memory, the
    // kernel rewrites the indices there in place, and they land in the record
    // at the context's next wait (every plan application waits before it
    // appends again, and presolve waits before it returns): one launch, no
    // synchronization of its own.
    // One staged block: the indices, then their kinds.
    const size_t bytes = sizeof(int32_t) * static_cast<size_t>(cnt);
    if (bytes + static_cast<size_t>(cnt) <= kStagedTapeMaxBytes) {
      std::vector<char> block(bytes + static_cast<size_t>(cnt));
      std::memcpy(block.data(), src.indices.data(), bytes);
      std::memcpy(block.data() + bytes, kind.data(), static_cast<size_t>(cnt));
      char* d_block = static_cast<char*>(
          c.stage(block.data(), block.size(), dst.indices.data() + base_idx, bytes, "tape indices"));
      check(gpk_map_tape_indices(static_cast<int32_t>(cnt), reinterpret_cast<int32_t*>(d_block),
                                 reinterpret_cast<const uint8_t*>(d_block + bytes), rs.row_r2o, rs.m, rs.col_r2o,
                                 rs.n, c.stream()),
            "tape globalize");
    } else {  // very long tape chunks go through device buffers
      DBuf<int32_t> d_idx(cnt);
      DBuf<uint8_t> d_kind(cnt);
      c.to_device(d_idx.get(), src.indices.data(), cnt, "tape indices upload");
      c.to_device(d_kind.get(), kind.data(), cnt, "tape kinds upload");
      check(gpk_map_tape_indices(static_cast<int32_t>(cnt), d_idx, d_kind, rs.row_r2o, rs.m, rs.col_r2o, rs.n,
                                 c.stream()),
            "tape globalize");
      c.download(dst.indices.data() + base_idx, d_idx.get(), bytes, "tape indices download");
      c.sync("tape indices");
    }
  }
  dst.types.insert(dst.types.end(), src.types.begin(), src.types.end());
  dst.dual_modes.insert(dst.dual_modes.end(), src.dual_modes.begin(), src.dual_modes.end());
  dst.vals.insert(dst.vals.end(), src.vals.begin(), src.vals.end());
  for (size_t k = 1; k < src.index_starts.size(); ++k) {
    dst.index_starts.push_back(base_idx + src.index_starts[k]);
    dst.value_starts.push_back(base_val + src.value_starts[k]);
  }
}

// update_record_from_plan (gpu_presolve.cu:1535-1802).
// fixed: the removed columns with a finite fixed interval (mask and its
// old-to-new map, k of them), computed by apply_plan with the keep maps.
struct FixedCols {
  DBuf<uint8_t> mask;
  DBuf<int32_t> map;
  int32_t k = 0;
};

// row_r2o_next / col_r2o_next: the maps already gathered through row_map /
// col_map (the structural path gathers them with the LP's vectors), or null.
void update_record(RecordState& rs, const PresolvePlanGpu& p, int32_t m, int32_t n, const int32_t* row_map,
                   const int32_t* col_map, int32_t m_new, int32_t n_new, const FixedCols& fixed,
                   DBuf<int32_t>* row_r2o_next = nullptr, DBuf<int32_t>* col_r2o_next = nullptr) {
  auto& c = ctx();
  hipStream_t s = c.stream();
  PresolveRecordGpu& rec = *rs.rec;
  if (fixed.k > 0) {
    // Removed columns with a finite fixed interval become (global index, midpoint).
    const int32_t k = fixed.k;
    DBuf<int32_t> oi(k);
    DBuf<double> ov(k);
    check(gpk_pack_fixed_records(n, fixed.mask, fixed.map, rs.col_r2o, p.new_l, p.new_u, oi, ov, s),
          "fixed records");
    const size_t base = rec.fixed_idx.size();
    rec.fixed_idx.resize(base + static_cast<size_t>(k));
    rec.fixed_val.resize(base + static_cast<size_t>(k));
    // Land at the context's next wait, like the tape indices below.
    c.begin_copy_group();
    c.download(rec.fixed_idx.data() + base, oi.get(), sizeof(int32_t) * k, "fixed records idx");
    c.download(rec.fixed_val.data() + base, ov.get(), sizeof(double) * k, "fixed records val");
    c.end_copy_group("fixed records");
  }
  append_global_tape(rs, p.tape);
  if (m_new < m) {
    if (row_r2o_next) {
      rs.row_r2o = std::move(*row_r2o_next);
    } else {
      DBuf<int32_t> next(m_new);
      check(gpk_gather_i32(row_map, m, rs.row_r2o, next, s), "row red2org");
      rs.row_r2o = std::move(next);
    }
    rs.m = m_new;
  }
  if (n_new < n) {
    if (col_r2o_next) {
      rs.col_r2o = std::move(*col_r2o_next);
    } else {
      DBuf<int32_t> next(n_new);
      check(gpk_gather_i32(col_map, n, rs.col_r2o, next, s), "col red2org");
      rs.col_r2o = std::move(next);
    }
    rs.n = n_new;
  }
  amd::record_update_checkpoint();
}

// apply_plan_to_working_lp (gpu_presolve.cu:2021-2147).
// new_at: the transpose of p.new_A (rows sorted) when the rule that rebuilt
// A formed it, else null.
void apply_plan(WorkLp& cur, const PresolvePlanGpu& p, RecordState& rs, const DeviceCsrMatrix* new_at = nullptr) {
  auto& c = ctx();
  hipStream_t s = c.stream();
  const int32_t m = cur.lp.A.rows, n = cur.lp.A.cols;
  // The row and column keep maps and the removed-fixed columns, with their
  // three counts read back in one synchronization.
  DBuf<int32_t> row_map(static_cast<int64_t>(m) + 1), col_map(static_cast<int64_t>(n) + 1);
  FixedCols fixed;
  fixed.mask.alloc(n);
  fixed.map.alloc(static_cast<int64_t>(n) + 1);
  int32_t* counts = c.flags();
  // Rank words of both maps (the column map's for compacting A, the row
  // map's for compacting AT) come out of the same launch.
  DBuf<gpk_rank_word> row_rank((static_cast<int64_t>(m) + 63) / 64), col_rank((static_cast<int64_t>(n) + 63) / 64);
  const gpk_apply_maps_args maps{m, n, p.keep_row_mask, p.keep_col_mask, p.new_l, p.new_u, row_map, col_map,
                                 fixed.mask, fixed.map, counts, row_rank.get(), col_rank.get()};
  check(gpk_apply_maps(maps, c.lookback(gpk_apply_maps_tiles(m, n)), s), "keep maps");
  // The caller's AT (first change) is compacted like a working AT when its
  // rows are sorted; the flag comes back with the counts.
  const bool callers_at = !cur.owns && !p.has_new_A;
  if (callers_at && cur.lp.AT.rows > 0 && cur.lp.AT.nnz > 1)
    check(gpk_rows_unsorted(cur.lp.AT.rowPtr, cur.lp.AT.colVal, cur.lp.AT.rows, counts + 3, s), "AT sorted check");
  // A plan that may drop rows or columns (or rebuilt A) has A's map pass
  // and tile scan launched before the sizes are read, so the kept entry
  // count comes back in the same wait as the three counts.
  const DeviceCsrMatrix& src = p.has_new_A ? p.new_A : cur.lp.A;
  const bool early = p.has_new_A || p.has_row_action || p.has_col_action;
  // When the AT to compact is known before the wait (the working AT, or the
  // rebuilding rule's transpose of its new A), both matrices' map passes go
  // out in one launch now and both write passes in one launch after it.
  const DeviceCsrMatrix* at_early = !early ? nullptr : p.has_new_A ? new_at : cur.owns ? &cur.lp.AT : nullptr;
  amd::CompactStart a_start;
  amd::CompactPairStart pair_start;
  if (at_early)
    pair_start =
        amd::compact_start_pair(src, row_map, col_rank.get(), *at_early, col_map, row_rank.get(), counts + 4);
  else if (early)
    a_start = amd::compact_start(src, row_map, col_rank.get(), counts + 4);
  const int32_t* h = c.read_flags(early ? 5 : 4);
  const bool callers_at_sorted = callers_at && h[3] == 0;
  const int32_t m_new = h[0], n_new = h[1];
  const int32_t a_nnz = early ? h[4] : -1;
  fixed.k = h[2];
  const double obj_new = cur.lp.obj_constant + p.obj_constant_delta;
  if (cur.owns && !p.has_new_A && m_new == m && n_new == n) {
    update_record(rs, p, m, n, row_map, col_map, m_new, n_new, fixed);
    amd::CopyBatch()
        .copy(cur.lp.c, p.new_c, n)
        .copy(cur.lp.l, p.new_l, n)
        .copy(cur.lp.u, p.new_u, n)
        .copy(cur.lp.AL, p.new_AL, m)
        .copy(cur.lp.AU, p.new_AU, m)
        .run("apply vectors");
    cur.lp.obj_constant = obj_new;
    return;
  }
  LPInfoGpu next;
  if (at_early) {
    std::tie(next.A, next.AT) = amd::compact_finish_pair(pair_start, m_new, n_new, a_nnz);
  } else {
    next.A = early ? amd::compact_finish(a_start, m_new, n_new, a_nnz)
                   : amd::compact_csr(src, row_map, col_map, m_new, n_new, -1, col_rank.get());
    // An AT with sorted rows that is the exact transpose of the A being
    // compacted (the working AT, the caller's AT when sorted, the rebuilding
    // rule's AT of a new A) is compacted with the maps swapped, which gives
    // the transpose of the compacted A; otherwise the compacted A is
    // re-transposed.
    const DeviceCsrMatrix* at_src = p.has_new_A ? new_at : (cur.owns || callers_at_sorted ? &cur.lp.AT : nullptr);
    if (at_src != nullptr) {
      next.AT = amd::compact_csr(*at_src, col_map, row_map, n_new, m_new, next.A.nnz, row_rank.get());
    } else {
      next.AT = amd::transpose_csr(next.A);
    }
  }
  next.c = pool_f64(n_new, "working c");
  next.l = pool_f64(n_new, "working l");
  next.u = pool_f64(n_new, "working u");
  next.AL = pool_f64(m_new, "working AL");
  next.AU = pool_f64(m_new, "working AU");
  // The record's reduced-to-original maps shrink with the LP; they are
  // gathered into new buffers, so the record update below still reads the
  // current ones.
  DBuf<int32_t> row_r2o_next, col_r2o_next;
  if (m_new < m) row_r2o_next.alloc(m_new);
  if (n_new < n) col_r2o_next.alloc(n_new);
  const gpk_apply_gathers_args ga{m, n, col_map, p.new_c, p.new_l, p.new_u, rs.col_r2o,
                                  next.c, next.l, next.u, n_new < n ? col_r2o_next.get() : nullptr,
                                  row_map, p.new_AL, p.new_AU, rs.row_r2o,
                                  next.AL, next.AU, m_new < m ? row_r2o_next.get() : nullptr};
  check(gpk_apply_gathers(ga, s), "gather vectors");
  update_record(rs, p, m, n, row_map, col_map, m_new, n_new, fixed, m_new < m ? &row_r2o_next : nullptr,
                n_new < n ? &col_r2o_next : nullptr);
  next.obj_constant = obj_new;
  cur.release();
  cur.lp = next;
  cur.owns = true;
}

// Copies the working LP into hipMalloc'ed arrays the caller releases with
// hipFree (free_gpu_presolve_reduced_lp); pool blocks for a serving job.
LPInfoGpu export_lp(const LPInfoGpu& w) {
  auto& c = ctx();
  hipStream_t s = c.stream();
  LPInfoGpu out = w;
  const bool pooled = amd::job_mode().pooled_outputs;
  amd::CopyBatch copies;
  auto take = [&](auto*& dst, const auto* src, int64_t count, const char* what) {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
    const size_t bytes = sizeof(T) * static_cast<size_t>(count > 0 ? count : 1);
    void* p = nullptr;
    if (pooled) p = amd::dev_alloc(bytes, s, what);
    else check(hipMalloc(&p, bytes), what);
    dst = static_cast<T*>(p);
    copies.copy(dst, src, count);
  };
  take(out.A.rowPtr, w.A.rowPtr, static_cast<int64_t>(w.A.rows) + 1, "reduced A rowPtr");
  take(out.A.colVal, w.A.colVal, w.A.nnz, "reduced A colVal");
  take(out.A.nzVal, w.A.nzVal, w.A.nnz, "reduced A nzVal");
  take(out.AT.rowPtr, w.AT.rowPtr, static_cast<int64_t>(w.AT.rows) + 1, "reduced AT rowPtr");
  take(out.AT.colVal, w.AT.colVal, w.AT.nnz, "reduced AT colVal");
  take(out.AT.nzVal, w.AT.nzVal, w.AT.nnz, "reduced AT nzVal");
  take(out.c, w.c, w.A.cols, "reduced c");
  take(out.l, w.l, w.A.cols, "reduced l");
  take(out.u, w.u, w.A.cols, "reduced u");
  take(out.AL, w.AL, w.A.rows, "reduced AL");
  take(out.AU, w.AU, w.A.rows, "reduced AU");
  copies.run("reduced LP export");
  c.sync("reduced LP export");
  return out;
}

void rebuild_org2red(std::vector<int32_t>& org2red, const std::vector<int32_t>& red2org, int32_t count) {
  org2red.assign(static_cast<size_t>(count > 0 ? count : 0), -1);
  for (size_t r = 0; r < red2org.size(); ++r) {
    const int32_t o = red2org[r];
    if (o >= 0 && o < count) org2red[static_cast<size_t>(o)] = static_cast<int32_t>(r);
  }
}

void validate_lp(const LPInfoGpu& lp) {
  if (lp.A.rows < 0 || lp.A.cols < 0 || lp.A.nnz < 0) throw std::runtime_error("LPInfoGpu: negative dimension");
  if (lp.AT.rows != lp.A.cols || lp.AT.cols != lp.A.rows || lp.AT.nnz != lp.A.nnz)
    throw std::runtime_error("LPInfoGpu: AT is not the transpose shape of A");
  if ((lp.A.cols > 0 && (!lp.c || !lp.l || !lp.u)) || (lp.A.rows > 0 && (!lp.AL || !lp.AU)))
    throw std::runtime_error("LPInfoGpu: missing vector");
  if (!lp.A.rowPtr || !lp.AT.rowPtr || (lp.A.nnz > 0 && (!lp.A.colVal || !lp.A.nzVal || !lp.AT.colVal || !lp.AT.nzVal)))
    throw std::runtime_error("LPInfoGpu: missing CSR array");
}

// Per-step wall time of one presolve, printed to stderr when the environment
// sets GP_PRESOLVE_PROFILE (the reference's rule profiling prints, e.g.
// gpu_presolve.cu:2797-2812, play the same role).  Each timed step
// synchronizes the stream first and last, so the figures are exact but the
// run is slower; without the variable nothing is synchronized or timed.
class StepProfile {
 public:
  StepProfile() : on_(std::getenv("GP_PRESOLVE_PROFILE") != nullptr) {}
  ~StepProfile() {
    if (!on_) return;
    std::fprintf(stderr, "presolve profile (ms, calls, host synchronizations):\n");
    for (const auto& kv : t_)
      std::fprintf(stderr, "  %-34s %9.3f %6d %7lld\n", kv.first.c_str(), 1e3 * kv.second.t, kv.second.calls,
                   static_cast<long long>(kv.second.syncs));
  }
  template <class F>
  void run(const char* name, F&& f) {
    if (!on_) {
      f();
      return;
    }
    ctx().sync("profile");
    const int64_t s0 = ctx().sync_count();
    const auto t0 = std::chrono::steady_clock::now();
    f();
    const int64_t s1 = ctx().sync_count();
    ctx().sync("profile");
    auto& e = t_[name];
    ++e.calls;
    e.syncs += s1 - s0;
    e.t += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

 private:
  struct Entry {
    int calls = 0;
    int64_t syncs = 0;
    double t = 0.0;
  };
  bool on_;
  std::map<std::string, Entry> t_;
};

// run_gpu_presolve_impl (gpu_presolve.cu:2162-3400).  The six
// structure-specific rules are not implemented: with
// enable_structure_specific_rules they are screened where the reference runs
// them, and an LP on which one would act is refused (UnsupportedStructure);
// on every other LP they do nothing, as in the reference.
// While the scheduler runs, rules leave their objective shifts to the next
// wait (amd::add_ordered_masked_sum).  A plan's shift is read only after a
// wait (the one that applies it); settle() covers a plan discarded without
// one.
struct DeferObjectiveSums {
  amd::Ctx& c;
  bool prev;
  DeferObjectiveSums() : c(amd::ctx()), prev(c.defer_objective_sums) { c.defer_objective_sums = true; }
  ~DeferObjectiveSums() { c.defer_objective_sums = prev; }
  DeferObjectiveSums(const DeferObjectiveSums&) = delete;
  DeferObjectiveSums& operator=(const DeferObjectiveSums&) = delete;
  void settle() {
    if (c.has_after_wait()) c.sync("objective shifts");
  }
};

GpuPresolveSummary run_impl(const LPInfoGpu& lp, const PresolveParams& prm, bool keep_reduced) {
  amd::AbandonOnThrow abandon_guard;
  DeferObjectiveSums defer_sums;
  validate_lp(lp);
  if (amd::job_mode().device_sync_on_entry) check(hipDeviceSynchronize(), "presolve entry synchronize");
  amd::check_lp_arrays(lp, "presolve input");
  GpuPresolveSummary res;
  res.original_rows = lp.A.rows;
  res.original_cols = lp.A.cols;
  PresolveRecordGpu& rec = res.record;
  rec.m0 = rec.m1 = lp.A.rows;
  rec.n0 = rec.n1 = lp.A.cols;
  rec.obj_constant_old = rec.obj_constant_new = lp.obj_constant;
  RecordState rs;
  rs.rec = &rec;
  rs.m = lp.A.rows;
  rs.n = lp.A.cols;
  rs.row_r2o.alloc(rs.m);
  rs.col_r2o.alloc(rs.n);
  if (rs.m > 0) check(gpk_iota_i32(rs.row_r2o, rs.m, ctx().stream()), "row red2org init");
  if (rs.n > 0) check(gpk_iota_i32(rs.col_r2o, rs.n, ctx().stream()), "col red2org init");

  WorkLp cur;
  cur.lp = lp;
  cur.owns = false;

  bool terminal = false;
  StepProfile prof;
  const auto t0 = std::chrono::steady_clock::now();
  const bool has_limit = prm.max_time > 0.0 && std::isfinite(prm.max_time);
  bool cs_dual_clean = false, cs_eq_clean = false;
  // The last phase's plan and statistics, kept while the LP is unchanged
  // and the plan pristine (plan_pristine).
  std::optional<PhasePlan> kept_plan;
  std::optional<PhaseStats> kept_stats;
  bool kept_valid = false;
  auto invalidate_clean = [&]() {
    cs_dual_clean = false;
    cs_eq_clean = false;
    kept_valid = false;  // the LP changed
  };
  auto phase_begin = [&]() -> PhasePlan& {
    if (kept_valid) {
      if (amd::plan_reuse_checked()) verify_kept_plan(kept_plan->p, kept_stats->s, cur.lp);
      amd::count_plan_reuse();
      return *kept_plan;
    }
    kept_stats.reset();
    kept_plan.reset();
    prof.run("plan + stats", [&] {
      kept_plan.emplace(cur.lp, false);
      kept_stats.emplace(cur.lp, &kept_plan->p);
    });
    return *kept_plan;
  };
  auto time_exceeded = [&]() {
    if (!has_limit) return false;
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >= prm.max_time;
  };
  // The doubleton rule's transpose of the A it rebuilt, for the plan's
  // application (freed after every plan).
  struct OwnedCsr {
    DeviceCsrMatrix m;
    ~OwnedCsr() { amd::free_csr(m); }
  } dbl_at;
  auto finish_plan = [&](PresolvePlanGpu& p) -> bool {
    res.has_infeasible = p.has_infeasible;
    res.has_unbounded = p.has_unbounded;
    terminal = p.has_infeasible || p.has_unbounded;
    const bool changed = !terminal && p.has_change;
    if (changed) {
      const DeviceCsrMatrix* at = p.has_new_A && dbl_at.m.rowPtr ? &dbl_at.m : nullptr;
      prof.run("apply plan", [&] { apply_plan(cur, p, rs, at); });
      invalidate_clean();
    }
    amd::free_csr(dbl_at.m);
    dbl_at.m = DeviceCsrMatrix{};
    if (!changed) defer_sums.settle();
    kept_valid = !changed && !terminal && plan_pristine(p);
    return changed;
  };

  // run_row_phase (gpu_presolve.cu:2219-2301)
  auto row_phase = [&](bool empty, bool singleton, bool redundant, bool implied, bool dup) -> bool {
    if (terminal || time_exceeded()) return false;
    PhasePlan& pp = phase_begin();
    PhaseStats& st = *kept_stats;
    PresolvePlanGpu& p = pp.p;
    const LPInfoGpu& v = cur.lp;
    if (!time_exceeded() && empty && prm.enable_empty_rows)
      prof.run("rule empty_rows", [&] { apply_rule_empty_rows(p, v, st.s, prm); });
    if (!time_exceeded() && singleton && prm.enable_singleton_rows)
      prof.run("rule singleton_rows", [&] { apply_rule_singleton_rows(p, v, st.s, prm); });
    if (!time_exceeded() && redundant && prm.enable_infeasible_redundant_rows)
      prof.run("rule infeasible_redundant_rows", [&] { apply_rule_infeasible_redundant_rows(p, v, st.s, prm); });
    if (!time_exceeded() && implied && prm.enable_implied_variable_bounds)
      prof.run("rule implied_variable_bounds", [&] { apply_rule_implied_variable_bounds(p, v, st.s, prm); });
    if (!time_exceeded() && dup && prm.enable_duplicate_rows)
      prof.run("rule duplicate_rows", [&] { apply_rule_duplicate_rows(p, v, st.s, prm); });
    return finish_plan(p);
  };

  // run_col_phase (gpu_presolve.cu:2303-2452)
  auto col_phase = [&](bool fixed, bool empty_cols, bool cs_dual, bool cs_eq, bool doubleton, bool dual_fix,
                       bool dup_cols, bool redundant_bounds) -> bool {
    if (terminal || time_exceeded()) return false;
    PhasePlan& pp = phase_begin();
    PhaseStats& st = *kept_stats;
    PresolvePlanGpu& p = pp.p;
    const LPInfoGpu& v = cur.lp;
    const bool run_redundant =
        redundant_bounds && (prm.enable_redundant_bounds || prm.enable_zero_cost_redundant_box_bounds);
    bool empty_done = false;
    if (!time_exceeded() && fixed && prm.enable_infeasible_fixed_variables) {
      if (empty_cols && prm.enable_empty_cols)
        prof.run("rule infeasible_fixed_variables",
                 [&] { empty_done = apply_rule_fixed_variables_and_empty_cols(p, v, st.s, prm); });
      else
        prof.run("rule infeasible_fixed_variables", [&] { apply_rule_infeasible_fixed_variables(p, v, prm); });
    }
    if (!empty_done && !time_exceeded() && empty_cols && prm.enable_empty_cols)
      prof.run("rule empty_cols", [&] { apply_rule_empty_cols(p, v, st.s, prm); });
    if (!time_exceeded() && cs_dual && prm.enable_column_singletons_dual_infer)
      prof.run("rule column_singletons_dual_infer", [&] { apply_rule_column_singletons_dual_infer(p, v, st.s, prm); });
    if (!time_exceeded() && cs_eq && prm.enable_column_singletons_eq)
      prof.run("rule column_singletons_eq", [&] { apply_rule_column_singletons_eq(p, v, st.s, prm, false); });
    if (!time_exceeded() && doubleton && prm.enable_doubleton_equations)
      prof.run("rule doubleton_equations", [&] {
        amd::free_csr(dbl_at.m);
        apply_rule_doubleton_equations_with_at(p, v, prm, &dbl_at.m);
      });
    if (!time_exceeded() && dual_fix && prm.enable_dual_fix)
      prof.run("rule dual_fix", [&] { apply_rule_dual_fix(p, v, st.s, prm, nullptr); });
    if (!time_exceeded() && dup_cols && prm.enable_duplicate_columns)
      prof.run("rule duplicate_columns", [&] { apply_rule_duplicate_columns(p, v, st.s, prm); });
    if (!time_exceeded() && run_redundant)
      prof.run("rule redundant_bounds", [&] { apply_rule_redundant_bounds(p, v, st.s, prm); });
    const bool changed = finish_plan(p);
    if (!changed) {
      const bool only_dual = !fixed && !empty_cols && cs_dual && !cs_eq && !doubleton && !dual_fix && !dup_cols &&
                             !redundant_bounds && prm.enable_column_singletons_dual_infer;
      const bool only_eq = !fixed && !empty_cols && !cs_dual && cs_eq && !doubleton && !dual_fix && !dup_cols &&
                           !redundant_bounds && prm.enable_column_singletons_eq;
      if (only_dual) cs_dual_clean = true;
      if (only_eq) cs_eq_clean = true;
    }
    return changed;
  };

  auto singleton_rows_to_exhaustion = [&]() {
    bool any = false;
    while (!terminal && !time_exceeded()) {
      const bool ch = row_phase(false, true, false, false, false);
      any = any || ch;
      if (!ch) break;
    }
    return any;
  };
  auto trivial_cleanup = [&]() {
    bool any = false;
    while (!terminal && !time_exceeded()) {
      bool pass = false;
      pass = col_phase(true, true, false, false, false, false, false, false) || pass;
      pass = row_phase(true, true, false, false, false) || pass;
      pass = col_phase(false, true, false, false, false, false, false, false) || pass;
      any = pass || any;
      if (!pass) break;
    }
    return any;
  };
  auto cleanup = [&]() {
    bool any = false;
    any = col_phase(true, true, false, false, false, false, false, false) || any;
    any = col_phase(false, false, false, false, false, true, false, false) || any;
    any = singleton_rows_to_exhaustion() || any;
    any = row_phase(true, false, false, false, false) || any;
    any = col_phase(false, true, false, false, false, false, false, false) || any;
    return any;
  };
  auto repeated_doubleton = [&]() {
    bool any = false;
    while (!terminal && !time_exceeded()) {
      const bool ch = col_phase(false, false, false, false, true, false, false, false);
      any = any || ch;
      if (!ch || prm.doubleton_equations_single_batch_per_iter) break;
    }
    return any;
  };
  // run_batched_singleton_cleanup_closure (gpu_presolve.cu:2503-2831):
  // column-singleton EQ, dual fix over incremental candidates and singleton
  // rows, repeated on one plan whose live degrees and singleton statistics
  // are refreshed from the masks on the device; one compaction at the end.
  auto singleton_closure = [&]() -> bool {
    if (terminal || time_exceeded() || cur.lp.A.nnz < amd::batched_closure_min_nnz() ||
        !prm.enable_column_singletons_eq || !prm.enable_dual_fix || !prm.enable_singleton_rows)
      return false;
    const LPInfoGpu& v = cur.lp;
    const int32_t m = v.A.rows, n = v.A.cols;
    hipStream_t s = ctx().stream();
    PhasePlan pp(v, false);
    PhaseStats st(v, &pp.p);
    PresolvePlanGpu& p = pp.p;
    const gpk_plan pv = amd::plan_view(p);
    const gpk_csr av = amd::csr_view(v.A), atv = amd::csr_view(v.AT);
    DBuf<uint8_t> row_before(m), col_before(n), cand(n), row_seen(m);
    DBuf<double> c_seen(n), l_seen(n), u_seen(n), AL_seen(m), AU_seen(m);
    if (m > 0) check(hipMemcpyAsync(row_before, p.keep_row_mask, m, hipMemcpyDeviceToDevice, s), "closure masks");
    if (n > 0) check(hipMemcpyAsync(col_before, p.keep_col_mask, n, hipMemcpyDeviceToDevice, s), "closure masks");
    const gpk_plan seen{row_seen, nullptr, c_seen, l_seen, u_seen, AL_seen, AU_seen};
    auto snapshot = [&]() { check(gpk_closure_snapshot(m, n, pv, seen, s), "closure snapshot"); };
    auto refresh = [&]() {
      check(gpk_closure_degrees(av, atv, p.keep_row_mask, p.keep_col_mask, row_before, col_before, st.s.row_nnz,
                                st.s.col_nnz, s),
            "closure degree refresh");
      check(gpk_closure_singletons(av, atv, p.keep_row_mask, p.keep_col_mask, st.s.row_nnz, st.s.col_nnz,
                                   st.s.singleton_row_mask, st.s.singleton_row_col, st.s.singleton_row_val,
                                   st.s.column_singleton_mask, st.s.column_singleton_row,
                                   st.s.column_singleton_val, s),
            "closure singleton refresh");
    };
    snapshot();
    bool any = false, first_dual = true;
    while (!terminal && !time_exceeded()) {
      p.has_change = false;
      apply_rule_column_singletons_eq(p, v, st.s, prm, true);
      const bool ch_eq = p.has_change;
      terminal = p.has_infeasible || p.has_unbounded;
      bool ch_dual = false;
      if (!terminal && !time_exceeded()) {
        if (first_dual) {
          if (n > 0) check(hipMemsetAsync(cand, 1, n, s), "closure first dual sweep");
          first_dual = false;
        } else {
          check(gpk_closure_candidates(av, pv, seen, cand, s), "closure dual candidates");
        }
        snapshot();
        p.has_change = false;
        apply_rule_dual_fix(p, v, st.s, prm, cand);
        ch_dual = p.has_change;
        terminal = p.has_infeasible || p.has_unbounded;
      }
      bool ch_rows = false;
      if (!terminal && !time_exceeded()) {
        refresh();
        p.has_change = false;
        apply_rule_singleton_rows(p, v, st.s, prm);
        ch_rows = p.has_change;
        terminal = p.has_infeasible || p.has_unbounded;
        if (!terminal && ch_rows) refresh();
      }
      const bool ch = ch_eq || ch_dual || ch_rows;
      any = any || ch;
      if (!ch) break;
    }
    res.has_infeasible = p.has_infeasible;
    res.has_unbounded = p.has_unbounded;
    p.has_change = any;
    if (any && !terminal) {
      apply_plan(cur, p, rs);
      invalidate_clean();
    }
    defer_sums.settle();  // before the closure's plan goes
    return any;
  };
  // run_fast_phase (gpu_presolve.cu:2833-2863)
  auto fast_phase = [&]() {
    bool any = singleton_closure();
    bool singleton = false;
    while (!terminal && !time_exceeded()) {
      const bool ch_dual = cs_dual_clean ? false : col_phase(false, false, true, false, false, false, false, false);
      const bool ch_eq = cs_eq_clean ? false : col_phase(false, false, false, true, false, false, false, false);
      const bool ch = ch_dual || ch_eq;
      any = any || ch;
      singleton = singleton || ch;
      if (ch) any = cleanup() || any;
      if (!ch) break;
    }
    const bool dbl = repeated_doubleton();
    any = dbl || any;
    if (dbl || singleton) any = cleanup() || any;
    return any;
  };
  // run_implied_variable_bounds_with_activity_closure (gpu_presolve.cu:2865-2987)
  auto implied_closure = [&]() {
    if (!prm.enable_implied_variable_bounds) return false;
    bool any = false;
    const int max_rounds = std::max(1, prm.implied_variable_bounds_max_rounds);
    const int cfg_bound_only = std::max(1, prm.implied_variable_bounds_max_bound_only_rounds);
    const int64_t pr = cur.lp.A.rows, pc = cur.lp.A.cols, pn = cur.lp.A.nnz;
    const bool wide_dense =
        pr >= 512 && pr <= 2048 && pc >= 128 * pr && pc <= 256 * pr && pn >= 2048 * pr && pn <= 8192 * pr;
    auto effective_limit = [&]() {
      int eff = cfg_bound_only;
      const int64_t budget = prm.implied_variable_bounds_bound_only_nnz_round_budget;
      if (budget <= 0 || cur.lp.A.nnz <= 0) return eff;
      const int64_t br = std::max<int64_t>(1, budget / cur.lp.A.nnz);
      if (br < eff) eff = static_cast<int>(br);
      if (wide_dense) eff = std::max(eff, std::min(2, cfg_bound_only));
      return eff;
    };
    int rounds = 0, bound_only = 0, limit = cfg_bound_only;
    while (rounds < max_rounds && !terminal && !time_exceeded()) {
      limit = effective_limit();
      PhasePlan pp(cur.lp, false);
      PhaseStats st(cur.lp, &pp.p);
      PresolvePlanGpu& p = pp.p;
      bool group_changed = false, structural = false;
      while (rounds < max_rounds && !terminal && !time_exceeded()) {
        p.has_change = false;
        apply_rule_implied_variable_bounds(p, cur.lp, st.s, prm);
        ++rounds;
        res.has_infeasible = p.has_infeasible;
        res.has_unbounded = p.has_unbounded;
        terminal = p.has_infeasible || p.has_unbounded;
        if (terminal || !p.has_change) break;
        group_changed = true;
        any = true;
        structural = p.has_new_A || p.has_row_action || p.has_col_action;
        if (structural) {
          bound_only = 0;
          break;
        }
        ++bound_only;
        if (bound_only >= limit) break;
      }
      if (group_changed && !terminal) {
        apply_plan(cur, p, rs);
        invalidate_clean();
      }
      if (terminal || !group_changed || !structural || bound_only >= limit) break;
    }
    if (any && !terminal && !time_exceeded()) any = row_phase(false, false, true, false, false) || any;
    return any;
  };
  auto medium_phase = [&]() {
    bool any = false, prop = false;
    prop = row_phase(false, false, true, false, false) || prop;
    prop = implied_closure() || prop;
    any = prop || any;
    if (prop) any = cleanup() || any;
    const bool dr = row_phase(false, false, false, false, true);
    any = dr || any;
    const bool dc = col_phase(false, false, false, false, false, false, true, false);
    any = dc || any;
    if (dr || dc) any = cleanup() || any;
    return any;
  };
  auto fixed_row_sequence = [&]() {
    bool any = false;
    any = row_phase(true, false, false, false, false) || any;
    any = row_phase(false, true, false, false, false) || any;
    any = row_phase(false, false, true, false, false) || any;
    const bool primal = implied_closure();
    any = primal || any;
    if (primal) any = trivial_cleanup() || any;
    any = row_phase(false, false, false, false, true) || any;
    return any;
  };
  auto trigger_col = [&](bool fixed, bool cs_dual, bool cs_eq, bool doubleton, bool dual_fix, bool dup) {
    const bool ch = col_phase(fixed, false, cs_dual, cs_eq, doubleton, dual_fix, dup, false);
    if (ch) (void)trivial_cleanup();
    return ch;
  };
  auto fixed_col_sequence = [&]() {
    bool any = false;
    any = trigger_col(true, false, false, false, false, false) || any;
    any = col_phase(false, true, false, false, false, false, false, false) || any;
    bool ch = false;
    do {
      ch = trigger_col(false, true, false, false, false, false);
      any = ch || any;
    } while (ch && !terminal && !time_exceeded());
    do {
      ch = trigger_col(false, false, true, false, false, false);
      any = ch || any;
    } while (ch && !terminal && !time_exceeded());
    do {
      ch = trigger_col(false, false, false, true, false, false);
      any = ch || any;
    } while (ch && !prm.doubleton_equations_single_batch_per_iter && !terminal && !time_exceeded());
    any = trigger_col(false, false, false, false, true, false) || any;
    any = trigger_col(false, false, false, false, false, true) || any;
    return any;
  };

  if (!prm.use_tiered_scheduler) {
    for (int iter = 0; iter < prm.max_iters && !terminal && !time_exceeded(); ++iter) {
      bool ch = false;
      ch = fixed_row_sequence() || ch;
      ch = fixed_col_sequence() || ch;
      ++res.iterations;
      if (!ch) break;
    }
  } else {
    if (prm.enable_tiered_bootstrap && prm.max_iters > 0 && !time_exceeded()) {
      // Bootstrap (gpu_presolve.cu:3082-3332).  The structure-specific rules
      // are screened where the reference runs them (the model-wide analyzers
      // and the structural L1 phase first, the structural L1 phase again
      // after a bootstrap that changed the LP); a screen that finds their
      // structure stops the presolve.
      amd::screen_bootstrap_structures(cur.lp, prm);
      bool generic = false;
      generic = row_phase(true, false, false, false, false) || generic;
      generic = row_phase(false, true, false, false, false) || generic;
      generic = row_phase(false, false, true, false, false) || generic;
      generic = row_phase(false, false, false, false, true) || generic;
      bool ch = col_phase(true, true, false, false, false, false, false, false);
      generic = ch || generic;
      if (ch) generic = trivial_cleanup() || generic;
      generic = col_phase(false, true, false, false, false, false, false, false) || generic;
      do {
        ch = col_phase(false, false, true, false, false, false, false, false);
        generic = ch || generic;
        if (ch) generic = trivial_cleanup() || generic;
      } while (ch && !terminal && !time_exceeded());
      do {
        ch = col_phase(false, false, false, true, false, false, false, false);
        generic = ch || generic;
        if (ch) generic = trivial_cleanup() || generic;
      } while (ch && !terminal && !time_exceeded());
      ch = col_phase(false, false, false, false, false, true, false, false);
      generic = ch || generic;
      if (ch) generic = trivial_cleanup() || generic;
      ch = col_phase(false, false, false, false, false, false, true, false);
      generic = ch || generic;
      if (ch) generic = trivial_cleanup() || generic;
      if (generic && !terminal && !time_exceeded()) amd::screen_structural_l1(cur.lp, prm);
      ++res.iterations;
    }
    bool fast = true;
    int32_t cycle_nnz_before = cur.lp.A.nnz;
    constexpr double kProgressRatio = 0.95;
    for (int iter = res.iterations; iter < prm.max_iters && !terminal && !time_exceeded(); ++iter) {
      bool ch = false;
      ch = cleanup() || ch;
      const int32_t nnz_before = cur.lp.A.nnz;
      if (fast) {
        ch = fast_phase() || ch;
        fast = _has_good_nnz_progress(nnz_before, cur.lp.A.nnz, kProgressRatio);
      } else {
        ch = medium_phase() || ch;
        const bool productive = _has_good_nnz_progress(cycle_nnz_before, cur.lp.A.nnz, kProgressRatio);
        if (!ch || !productive) {
          ++res.iterations;
          break;
        }
        cycle_nnz_before = cur.lp.A.nnz;
        fast = true;
      }
      ++res.iterations;
      if (!ch && fast) break;
    }
  }
  if (!terminal && !time_exceeded() && (prm.enable_redundant_bounds || prm.enable_zero_cost_redundant_box_bounds))
    (void)col_phase(false, false, false, false, false, false, false, true);

  auto& c = ctx();
  rec.row_red2org = c.to_host(rs.row_r2o.get(), rs.m, "record row red2org");
  rec.col_red2org = c.to_host(rs.col_r2o.get(), rs.n, "record col red2org");
  res.reduced_rows = cur.lp.A.rows;
  res.reduced_cols = cur.lp.A.cols;
  res.reduced_nnz = cur.lp.A.nnz;
  defer_sums.settle();
  res.obj_constant_delta = cur.lp.obj_constant - lp.obj_constant;
  rec.m1 = res.reduced_rows;
  rec.n1 = res.reduced_cols;
  rec.obj_constant_new = cur.lp.obj_constant;
  rebuild_org2red(rec.row_org2red, rec.row_red2org, rec.m0);
  rebuild_org2red(rec.col_org2red, rec.col_red2org, rec.n0);
  if (keep_reduced) {
    if (cur.owns) {
      res.reduced_lp = export_lp(cur.lp);
      res.owns_reduced_lp = true;
    } else {
      res.reduced_lp = cur.lp;  // unchanged: the caller's own LP (owns = false)
      res.owns_reduced_lp = false;
    }
  }
  cur.release();
  c.sync("presolve exit");
  res.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res;
}

}  // namespace

bool _has_good_nnz_progress(std::int32_t nnz_before, std::int32_t nnz_after, double ratio) {
  if (nnz_before <= 0) return false;
  return static_cast<double>(nnz_after) < ratio * static_cast<double>(nnz_before);
}

GpuPresolveSummary run_gpu_presolve_with_record(const LPInfoGpu& lp, const PresolveParams& params) {
  return run_impl(lp, params, false);
}

GpuPresolveSummary run_gpu_presolve_with_reduced_lp(const LPInfoGpu& lp, const PresolveParams& params) {
  return run_impl(lp, params, true);
}

GpuPresolveSummary::GpuPresolveSummary(GpuPresolveSummary&& o) noexcept
    : original_rows(std::exchange(o.original_rows, 0)),
      original_cols(std::exchange(o.original_cols, 0)),
      reduced_rows(std::exchange(o.reduced_rows, 0)),
      reduced_cols(std::exchange(o.reduced_cols, 0)),
      reduced_nnz(std::exchange(o.reduced_nnz, 0)),
      iterations(std::exchange(o.iterations, 0)),
      elapsed_seconds(std::exchange(o.elapsed_seconds, 0.0)),
      obj_constant_delta(std::exchange(o.obj_constant_delta, 0.0)),
      has_infeasible(std::exchange(o.has_infeasible, false)),
      has_unbounded(std::exchange(o.has_unbounded, false)),
      record(std::move(o.record)),
      reduced_lp(std::exchange(o.reduced_lp, LPInfoGpu{})),
      owns_reduced_lp(std::exchange(o.owns_reduced_lp, false)) {}

GpuPresolveSummary& GpuPresolveSummary::operator=(GpuPresolveSummary&& o) noexcept {
  if (this != &o) {
    free_gpu_presolve_reduced_lp(*this);
    original_rows = std::exchange(o.original_rows, 0);
    original_cols = std::exchange(o.original_cols, 0);
    reduced_rows = std::exchange(o.reduced_rows, 0);
    reduced_cols = std::exchange(o.reduced_cols, 0);
    reduced_nnz = std::exchange(o.reduced_nnz, 0);
    iterations = std::exchange(o.iterations, 0);
    elapsed_seconds = std::exchange(o.elapsed_seconds, 0.0);
    obj_constant_delta = std::exchange(o.obj_constant_delta, 0.0);
    has_infeasible = std::exchange(o.has_infeasible, false);
    has_unbounded = std::exchange(o.has_unbounded, false);
    record = std::move(o.record);
    reduced_lp = std::exchange(o.reduced_lp, LPInfoGpu{});
    owns_reduced_lp = std::exchange(o.owns_reduced_lp, false);
  }
  return *this;
}

void free_gpu_presolve_reduced_lp(GpuPresolveSummary& summary) {
  if (!summary.owns_reduced_lp) {
    summary.reduced_lp = LPInfoGpu{};
    return;
  }
  LPInfoGpu& r = summary.reduced_lp;
  for (void* p : {static_cast<void*>(r.A.rowPtr), static_cast<void*>(r.A.colVal), static_cast<void*>(r.A.nzVal),
                  static_cast<void*>(r.AT.rowPtr), static_cast<void*>(r.AT.colVal), static_cast<void*>(r.AT.nzVal),
                  static_cast<void*>(r.c), static_cast<void*>(r.AL), static_cast<void*>(r.AU),
                  static_cast<void*>(r.l), static_cast<void*>(r.u)})
    amd::dev_free(p, ctx().stream());  // hipFree, or back to the pool for a serving job's export
  summary.reduced_lp = LPInfoGpu{};
  summary.owns_reduced_lp = false;
}

// ---- PostsolveTapeGpu (presolve_structs.cu:9-60) ---------------------------
PostsolveTapeGpu::~PostsolveTapeGpu() { reset(); }

PostsolveTapeGpu::PostsolveTapeGpu(PostsolveTapeGpu&& o) noexcept
    : types(std::exchange(o.types, nullptr)),
      index_starts(std::exchange(o.index_starts, nullptr)),
      value_starts(std::exchange(o.value_starts, nullptr)),
      dual_modes(std::exchange(o.dual_modes, nullptr)),
      indices(std::exchange(o.indices, nullptr)),
      vals(std::exchange(o.vals, nullptr)),
      record_count(std::exchange(o.record_count, 0)),
      index_count(std::exchange(o.index_count, 0)),
      value_count(std::exchange(o.value_count, 0)),
      record_capacity(std::exchange(o.record_capacity, 0)),
      index_capacity(std::exchange(o.index_capacity, 0)),
      value_capacity(std::exchange(o.value_capacity, 0)) {}

PostsolveTapeGpu& PostsolveTapeGpu::operator=(PostsolveTapeGpu&& o) noexcept {
  if (this != &o) {
    reset();
    types = std::exchange(o.types, nullptr);
    index_starts = std::exchange(o.index_starts, nullptr);
    value_starts = std::exchange(o.value_starts, nullptr);
    dual_modes = std::exchange(o.dual_modes, nullptr);
    indices = std::exchange(o.indices, nullptr);
    vals = std::exchange(o.vals, nullptr);
    record_count = std::exchange(o.record_count, 0);
    index_count = std::exchange(o.index_count, 0);
    value_count = std::exchange(o.value_count, 0);
    record_capacity = std::exchange(o.record_capacity, 0);
    index_capacity = std::exchange(o.index_capacity, 0);
    value_capacity = std::exchange(o.value_capacity, 0);
  }
  return *this;
}

void PostsolveTapeGpu::reset() noexcept {
  for (void* p : {static_cast<void*>(types), static_cast<void*>(index_starts), static_cast<void*>(value_starts),
                  static_cast<void*>(dual_modes), static_cast<void*>(indices), static_cast<void*>(vals)})
    if (p) (void)hipFree(p);
  types = nullptr;
  index_starts = nullptr;
  value_starts = nullptr;
  dual_modes = nullptr;
  indices = nullptr;
  vals = nullptr;
  record_count = index_count = value_count = 0;
  record_capacity = index_capacity = value_capacity = 0;
}

}  // namespace gpu_presolver::presolve
