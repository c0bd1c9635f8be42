// Postsolve driver (reference cpp/src/presolve/gpu_postsolve.cu:1620-1721):
// scatter of the reduced solution, fixed-value restore, reverse tape replay
// (primal pass, then dual pass) and the final column-dual projection.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
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
  const size_t bytes = sizeof(double) * static_cast<size_t>(n);
  if (amd::job_mode().pooled_outputs) return static_cast<double*>(amd::dev_alloc(bytes, ctx().stream(), what));
  void* p = nullptr;
  check(hipMalloc(&p, bytes), what);
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

// ---- dependency-level replay schedule ---------------------------------------
// The reference replays the tape in reverse with one lane.  Here each pass is
// scheduled in levels: record k gets level 1 + the highest level among the
// later records (replayed before it) that write a variable it reads or
// writes, or read a variable it writes.  Records of one level touch disjoint
// variables and run in parallel; every variable sees its reads and writes in
// the serial order, so the results are the serial replay's, bit for bit.
// Variables: column c (x, z and z_retrieved) -> c, row r (y) -> n0 + r.

// Index lists of the original rows of A (columns of AT) that the dual
// recoveries walk, gathered on the device and copied to the host: line l is
// idx[off[slot[l]] .. off[slot[l] + 1]).
struct Lines {
  std::vector<int32_t> slot, off, idx;
  const int32_t* begin(int32_t l) const { return idx.data() + off[slot[l]]; }
  const int32_t* end(int32_t l) const { return idx.data() + off[slot[l] + 1]; }
};

struct OrigPattern {
  Lines a_rows, at_cols;
};

Lines gather_lines(const int32_t* ptr, const int32_t* idx, int32_t lines_total, const std::vector<int32_t>& wanted) {
  Lines L;
  L.slot.assign(static_cast<size_t>(lines_total), -1);
  std::vector<int32_t> list;
  for (int32_t l : wanted)
    if (L.slot[l] < 0) {
      L.slot[l] = static_cast<int32_t>(list.size());
      list.push_back(l);
    }
  const int32_t k = static_cast<int32_t>(list.size());
  L.off.assign(1, 0);
  if (k == 0) return L;
  auto& c = ctx();
  DBuf<int32_t> d_list(k), len(static_cast<int64_t>(k) + 1), off(static_cast<int64_t>(k) + 1);
  c.to_device(d_list.get(), list.data(), k, "pattern line list");
  check(gpk_seg_lengths(ptr, d_list, k, len, c.stream()), "pattern line lengths");
  const int32_t total = amd::scan_counts(len, k, off);
  DBuf<int32_t> out(total);
  check(gpk_seg_copy(ptr, idx, d_list, k, off, out, c.stream()), "pattern line copy");
  L.off = c.to_host(off.get(), static_cast<int64_t>(k) + 1, "pattern line offsets");
  L.idx = c.to_host(out.get(), total, "pattern lines");
  return L;
}

// Levels [first, first + count): one grid launch (a single large level) or
// one workgroup that steps through a run of smaller levels.
struct Segment {
  int32_t first, count;
  bool grid;
};

struct Schedule {
  std::vector<int32_t> order;        // record indices, level by level
  std::vector<int32_t> level_start;  // levels + 1 offsets into order
  std::vector<Segment> segs;
};

constexpr int32_t kGridLevelMin = 2048;

[[noreturn]] void bad_tape(int32_t k, const char* what) {
  throw std::runtime_error("postsolve_gpu: tape record " + std::to_string(k) + ": " + what);
}

class TapeScanner {
 public:
  TapeScanner(const PostsolveTape& t, int32_t n0, int32_t m0) : t_(t), n0_(n0), m0_(m0) {}

  // Read and write sets of record k in the primal (primal = true) or dual
  // pass; returns false when the record does nothing in that pass.
  bool sets(int32_t k, bool primal, bool has_orig, bool has_y, const OrigPattern* op, std::vector<int32_t>& rd,
            std::vector<int32_t>& wr) const {
    rd.clear();
    wr.clear();
    const int32_t type = t_.types[k];
    const int32_t i0 = t_.index_starts[k], i1 = t_.index_starts[k + 1] - 1;
    const int32_t v0 = t_.value_starts[k], v1 = t_.value_starts[k + 1] - 1;
    const int32_t* I = t_.indices.data();
    const double* V = t_.vals.data();
    const int32_t ni = i1 - i0 + 1, nv = v1 - v0 + 1;
    auto need = [&](bool ok) {
      if (!ok) bad_tape(k, "too few indices or values for its type");
    };
    switch (type) {
      case 3: {  // DuplicateColumn
        need(ni >= 2 && nv >= 5);
        const int32_t from = col(k, I[i0]), to = col(k, I[i0 + 1]);
        rd.push_back(to);
        wr.push_back(from);
        if (primal) wr.push_back(to);
        return true;
      }
      case 13: {  // DoubletonEquation
        need(ni >= 3 && nv >= 10);
        const int32_t ec = col(k, I[i0]), kc = col(k, I[i0 + 1]), row = I[i0 + 2];
        if (primal) {
          rd.push_back(kc);
          wr.push_back(ec);
          return true;
        }
        if (has_orig && has_y) {
          const int32_t rv = row_var(k, row);
          rd.push_back(kc);
          const bool payload = ni >= 4 && nv >= 12;
          if (payload) {
            need(ni >= 4 + I[i0 + 3] && nv >= 12 + I[i0 + 3]);
            for (int32_t t = 0; t < I[i0 + 3]; ++t) rd.push_back(row_var(k, I[i0 + 4 + t]));
          } else {
            for (const int32_t* q = op->at_cols.begin(ec); q != op->at_cols.end(ec); ++q)
              if (*q != row) rd.push_back(row_var(k, *q));
          }
          wr.push_back(rv);
          wr.push_back(kc);
          wr.push_back(ec);
        } else {
          wr.push_back(ec);
          if (has_y && row >= 0) wr.push_back(row_var(k, row));
        }
        return true;
      }
      case 2: {  // SubCol
        need(ni >= 3 && nv >= 6);
        const int32_t ec = col(k, I[i0]), row = I[i0 + 1], sc = I[i0 + 2];
        need(sc >= 0 && ni >= 3 + sc && nv >= 6 + sc);
        for (int32_t t = 0; t < sc; ++t) rd.push_back(col(k, I[i0 + 3 + t]));
        wr.push_back(ec);
        if (!primal && has_y && row >= 0) {
          rd.push_back(row_var(k, row));
          wr.push_back(row_var(k, row));
        }
        return true;
      }
      case 0: {  // FixedCol
        need(ni >= 1 && nv >= 1 + (primal ? 0 : 1) + (primal ? 0 : ni - 1));
        wr.push_back(col(k, I[i0]));
        if (!primal && has_y)
          for (int32_t pos = i0 + 1; pos <= i1; ++pos) rd.push_back(row_var(k, I[pos]));
        return true;
      }
      case 1: {  // FixedColInf
        need(ni >= 2 && nv >= 2);
        const int32_t c = col(k, I[i0 + 1]);
        wr.push_back(c);
        if (primal) {
          const int64_t nrows = std::llround(V[v0]);
          int32_t ip = i0 + 2, vp = v0 + 2;
          for (int64_t r = 0; r < nrows; ++r) {
            need(ip <= i1);
            const int32_t len = I[ip];
            need(len >= 0 && ip + len <= i1 && vp + len <= v1);
            for (int32_t q = 0; q < len; ++q) rd.push_back(col(k, I[ip + 1 + q]));
            ip += len + 1;
            vp += len + 1;
          }
        }
        return true;
      }
      case 14: {  // FmeCol
        need(ni >= 3 && nv >= 1);
        const int32_t K = I[i0];
        need(K >= 0 && ni >= 3 + K);
        const int32_t ec = col(k, I[i0 + 1]);
        wr.push_back(ec);
        if (primal) {
          const int64_t L = std::llround(V[v0]);
          int32_t cc = i0 + 3 + K, vp = v0 + 1;
          for (int64_t s = 0; s < L; ++s) {
            need(vp + 2 <= v1);
            const int64_t len = std::llround(V[vp + 1]);
            need(len >= 0 && cc + len - 1 <= i1 && vp + 2 + len <= v1);
            for (int64_t q = 0; q < len; ++q) rd.push_back(col(k, I[cc++]));
            vp += 3 + static_cast<int32_t>(len);
          }
        } else if (has_y) {
          for (int32_t q = 0; q < K; ++q) wr.push_back(row_var(k, I[i0 + 2 + q]));
        }
        return true;
      }
      case 10:  // EqToIneq
        if (primal || !has_y) return false;
        need(ni >= 1 && nv >= 1);
        rd.push_back(row_var(k, I[i0]));
        wr.push_back(row_var(k, I[i0]));
        return true;
      case 4:  // DuplicateRow
        if (primal || !has_y || !(ni >= 2 && nv >= 5)) return false;
        for (int32_t q = 0; q < 2; ++q) {
          rd.push_back(row_var(k, I[i0 + q]));
          wr.push_back(row_var(k, I[i0 + q]));
        }
        return true;
      case 5: {  // DeletedRow
        if (primal || !has_orig || !(ni >= 2 && nv >= 3)) return false;
        const int32_t row = I[i0], c = col(k, I[i0 + 1]);
        rd.push_back(c);
        if (has_y) {
          for (const int32_t* q = op->at_cols.begin(c); q != op->at_cols.end(c); ++q)
            if (*q != row) rd.push_back(row_var(k, *q));
          wr.push_back(row_var(k, row));
        }
        wr.push_back(c);
        return true;
      }
      case 12: {  // BoundChangeTheRow
        if (primal || !has_orig || ni < 2) return false;
        need(nv >= 4);
        const int32_t c = col(k, I[i0]), row = I[i0 + 1];
        const int32_t rv = row_var(k, row);
        rd.push_back(c);
        wr.push_back(c);
        if (has_y) {
          rd.push_back(rv);
          wr.push_back(rv);
        }
        const int32_t r = rv - n0_;
        for (const int32_t* q = op->a_rows.begin(r); q != op->a_rows.end(r); ++q) {
          rd.push_back(col(k, *q));
          wr.push_back(*q);
        }
        return true;
      }
      default:
        return false;
    }
  }

  // Original rows (type 12) and columns (types 5 and 13 without a row
  // payload) whose index lists the dual pass walks.
  void pattern_lines(std::vector<int32_t>& rows, std::vector<int32_t>& cols) const {
    for (int32_t k = 0; k < static_cast<int32_t>(t_.types.size()); ++k) {
      const int32_t type = t_.types[k];
      const int32_t i0 = t_.index_starts[k];
      const int32_t ni = t_.index_starts[k + 1] - i0;
      const int32_t nv = t_.value_starts[k + 1] - t_.value_starts[k];
      if (type == 12 && ni >= 2) {
        row_var(k, t_.indices[i0 + 1]);
        rows.push_back(t_.indices[i0 + 1]);
      } else if (type == 5 && ni >= 2) {
        cols.push_back(col(k, t_.indices[i0 + 1]));
      } else if (type == 13 && ni >= 3 && !(ni >= 4 && nv >= 12)) {
        cols.push_back(col(k, t_.indices[i0]));
      }
    }
  }

 private:
  int32_t col(int32_t k, int32_t c) const {
    if (c < 0 || c >= n0_) bad_tape(k, "column index out of range");
    return c;
  }
  int32_t row_var(int32_t k, int32_t r) const {
    if (r < 0 || r >= m0_) bad_tape(k, "row index out of range");
    return n0_ + r;
  }
  const PostsolveTape& t_;
  int32_t n0_, m0_;
};

// Records [lo, hi) of the tape, replayed in reverse.
Schedule level_schedule(const TapeScanner& scan, int32_t lo, int32_t hi, int32_t n0, int32_t m0, bool primal,
                        bool has_orig, bool has_y, const OrigPattern* op) {
  std::vector<int32_t> lw(static_cast<size_t>(n0) + m0, -1), lr(static_cast<size_t>(n0) + m0, -1);
  const int32_t span = hi > lo ? hi - lo : 0;
  std::vector<int32_t> level(span, -1);
  std::vector<int32_t> rd, wr;
  int32_t levels = 0;
  for (int32_t k = hi - 1; k >= lo; --k) {
    if (!scan.sets(k, primal, has_orig, has_y, op, rd, wr)) continue;
    int32_t lv = 0;
    for (int32_t v : rd) lv = std::max(lv, lw[v] + 1);
    for (int32_t v : wr) lv = std::max(lv, std::max(lw[v], lr[v]) + 1);
    for (int32_t v : wr) lw[v] = lv;
    for (int32_t v : rd) lr[v] = std::max(lr[v], lv);
    level[k - lo] = lv;
    levels = std::max(levels, lv + 1);
  }
  std::vector<int32_t> start(static_cast<size_t>(levels) + 1, 0);
  for (int32_t k = 0; k < span; ++k)
    if (level[k] >= 0) ++start[level[k] + 1];
  for (int32_t l = 0; l < levels; ++l) start[l + 1] += start[l];
  Schedule sch;
  sch.order.resize(start[levels]);
  std::vector<int32_t> fill(start.begin(), start.end() - 1);
  for (int32_t k = hi - 1; k >= lo; --k)
    if (level[k - lo] >= 0) sch.order[fill[level[k - lo]]++] = k;  // descending k within a level
  for (int32_t l = 0; l < levels; ++l) {
    const bool big = start[l + 1] - start[l] >= kGridLevelMin;
    if (!big && !sch.segs.empty() && !sch.segs.back().grid) {
      ++sch.segs.back().count;
    } else {
      sch.segs.push_back(Segment{l, 1, big});
    }
  }
  sch.level_start = std::move(start);
  return sch;
}

// The schedule's record order and level bounds in one device block.
DBuf<int32_t> upload_schedule(const Schedule& sch) {
  const int64_t nr = static_cast<int64_t>(sch.order.size()), nl = static_cast<int64_t>(sch.level_start.size());
  DBuf<int32_t> buf(nr + nl);
  amd::Ctx& c = ctx();
  c.begin_copy_group();
  c.to_device(buf.get(), sch.order.data(), nr, "replay schedule upload");
  c.to_device(buf.get() + nr, sch.level_start.data(), nl, "replay schedule upload");
  c.end_copy_group("replay schedule upload");
  return buf;
}

// Launches only: no allocation, copy or wait, so it can be captured.
void enqueue_schedule(const Schedule& sch, const int32_t* buf, const gp_tape_dev& tape, const gp_orig_dev& orig,
                      double* x, double* y, double* z, uint8_t* zret, bool primal, double tol) {
  if (sch.order.empty()) return;
  hipStream_t s = ctx().stream();
  const int32_t* order = buf;
  const int32_t* starts = buf + sch.order.size();
  const int32_t p = primal ? 1 : 0;
  const char* what = primal ? "replay primal" : "replay dual";
  for (const Segment& g : sch.segs) {
    if (g.grid) {
      const int32_t b = sch.level_start[g.first];
      check(gpk_replay_list(tape, orig, x, y, z, zret, order + b, sch.level_start[g.first + 1] - b, p, 1 - p, tol, s),
            what);
    } else {
      check(gpk_replay_levels(tape, orig, x, y, z, zret, order, starts + g.first, g.count, p, 1 - p, tol, s), what);
    }
  }
}

void upload_map(DBuf<int32_t>& dst, const std::vector<int32_t>& src, int32_t bound, const char* what) {
  for (int32_t v : src)
    if (v < 0 || v >= bound) throw std::runtime_error(std::string(what) + ": index out of range");
  dst.alloc(static_cast<int64_t>(src.size()));
  ctx().to_device(dst.get(), src.data(), static_cast<int64_t>(src.size()), what);
}

}  // namespace

}  // namespace gpu_presolver::presolve

namespace gpu_presolver::amd {

using presolve::kPostsolveTol;
using presolve::LPInfoGpu;
using presolve::PresolveRecordGpu;

namespace {

void check_recovery_col(int32_t col, int32_t n0, const char* what) {
  if (col < 0 || col >= n0) throw std::runtime_error(std::string(what) + " column is out of range");
}

// validate_primal_recovery_timeline (gpu_postsolve.cu:1485-1616): the
// checkpoints are chronological, each payload has exactly one checkpoint
// and every column a payload names is an original column.
void validate_recovery_timeline(const PresolveRecordGpu& r) {
  const auto limit = static_cast<size_t>(std::numeric_limits<int32_t>::max());
  if (r.structural_primal_recoveries.size() > limit || r.antipodal_component_recoveries.size() > limit)
    throw std::runtime_error("primal recovery payload exceeds 32-bit checkpoint range");
  std::vector<uint8_t> seen_s(r.structural_primal_recoveries.size(), 0), seen_a(r.antipodal_component_recoveries.size(), 0);
  const int32_t tape_count = static_cast<int32_t>(r.tape.types.size());
  int32_t prev = 0;
  for (const presolve::PrimalRecoveryCheckpoint& cp : r.primal_recovery_timeline) {
    if (cp.tape_position < prev || cp.tape_position > tape_count)
      throw std::runtime_error("primal recovery checkpoints are not chronological");
    prev = cp.tape_position;
    if (cp.kind == presolve::PrimalRecoveryKind::StructuralL1) {
      if (cp.payload_index < 0 || cp.payload_index >= static_cast<int32_t>(seen_s.size()) || seen_s[cp.payload_index])
        throw std::runtime_error("invalid or duplicate structural recovery checkpoint");
      seen_s[cp.payload_index] = 1;
      const auto& st = r.structural_primal_recoveries[cp.payload_index];
      for (const auto& x : st.splits) {
        check_recovery_col(x.t_col, r.n0, "structural split recovery");
        check_recovery_col(x.e_col, r.n0, "structural split recovery");
      }
      for (const auto& x : st.outer_pairs) {
        check_recovery_col(x.bound_col, r.n0, "structural outer-pair recovery");
        check_recovery_col(x.free_col, r.n0, "structural outer-pair recovery");
      }
      for (const auto& x : st.linked_slacks) {
        check_recovery_col(x.slack_col, r.n0, "structural linked-slack recovery");
        check_recovery_col(x.t_col, r.n0, "structural linked-slack recovery");
      }
      for (const auto& x : st.max_slacks) {
        if (x.t_cols.size() != x.factors.size())
          throw std::runtime_error("structural max-slack recovery has mismatched payloads");
        check_recovery_col(x.slack_col, r.n0, "structural max-slack recovery");
        for (int32_t col : x.t_cols) check_recovery_col(col, r.n0, "structural max-slack recovery");
      }
    } else if (cp.kind == presolve::PrimalRecoveryKind::AntipodalComponent) {
      if (cp.payload_index < 0 || cp.payload_index >= static_cast<int32_t>(seen_a.size()) || seen_a[cp.payload_index])
        throw std::runtime_error("invalid or duplicate antipodal recovery checkpoint");
      seen_a[cp.payload_index] = 1;
      const auto& a = r.antipodal_component_recoveries[cp.payload_index];
      const size_t k = a.plus_cols.size();
      if (a.minus_cols.size() != k || a.root_plus_cols.size() != k || a.root_minus_cols.size() != k ||
          a.plus_lower.size() != k || a.minus_lower.size() != k || k > limit)
        throw std::runtime_error("antipodal recovery has mismatched payloads");
      for (size_t i = 0; i < k; ++i) {
        check_recovery_col(a.plus_cols[i], r.n0, "antipodal recovery");
        check_recovery_col(a.minus_cols[i], r.n0, "antipodal recovery");
        check_recovery_col(a.root_plus_cols[i], r.n0, "antipodal recovery");
        check_recovery_col(a.root_minus_cols[i], r.n0, "antipodal recovery");
      }
    } else {
      throw std::runtime_error("unknown primal recovery checkpoint kind");
    }
  }
  for (uint8_t v : seen_s)
    if (!v) throw std::runtime_error("structural recovery payload is missing a checkpoint");
  for (uint8_t v : seen_a)
    if (!v) throw std::runtime_error("antipodal recovery payload is missing a checkpoint");
}

// One recovery step on the device: its payload arrays uploaded once.
struct RecoveryStep {
  bool antipodal = false;
  int32_t nsplit = 0, nouter = 0, nlinked = 0, nmax = 0, nanti = 0;
  DBuf<int32_t> split_t, split_e, outer_b, outer_f, linked_s, linked_t, max_s, max_start, max_t;
  DBuf<double> split_rho, linked_f, max_f;
  DBuf<int32_t> a_plus, a_minus, a_rplus, a_rminus;
  DBuf<double> a_plower, a_mlower, a_pcache, a_mcache;

  template <class T>
  static void up(DBuf<T>& d, const std::vector<T>& v, const char* what) {
    d.alloc(static_cast<int64_t>(v.size()));
    ctx().to_device(d.get(), v.data(), static_cast<int64_t>(v.size()), what);
  }

  explicit RecoveryStep(const presolve::StructuralL1PrimalRecoveryStep& st) {
    std::vector<int32_t> t, e, b, f, ls, lt, ms, mstart{0}, mt;
    std::vector<double> rho, lf, mf;
    for (const auto& x : st.splits) {
      t.push_back(x.t_col);
      e.push_back(x.e_col);
      rho.push_back(x.rho);
    }
    for (const auto& x : st.outer_pairs) {
      b.push_back(x.bound_col);
      f.push_back(x.free_col);
    }
    for (const auto& x : st.linked_slacks) {
      ls.push_back(x.slack_col);
      lt.push_back(x.t_col);
      lf.push_back(x.factor);
    }
    for (const auto& x : st.max_slacks) {
      if (x.t_cols.empty()) continue;  // the reference skips an empty max slack
      ms.push_back(x.slack_col);
      mt.insert(mt.end(), x.t_cols.begin(), x.t_cols.end());
      mf.insert(mf.end(), x.factors.begin(), x.factors.end());
      mstart.push_back(static_cast<int32_t>(mt.size()));
    }
    nsplit = static_cast<int32_t>(t.size());
    nouter = static_cast<int32_t>(b.size());
    nlinked = static_cast<int32_t>(ls.size());
    nmax = static_cast<int32_t>(ms.size());
    up(split_t, t, "split recovery upload");
    up(split_e, e, "split recovery upload");
    up(split_rho, rho, "split recovery upload");
    up(outer_b, b, "outer-pair recovery upload");
    up(outer_f, f, "outer-pair recovery upload");
    up(linked_s, ls, "linked-slack recovery upload");
    up(linked_t, lt, "linked-slack recovery upload");
    up(linked_f, lf, "linked-slack recovery upload");
    up(max_s, ms, "max-slack recovery upload");
    up(max_start, mstart, "max-slack recovery upload");
    up(max_t, mt, "max-slack recovery upload");
    up(max_f, mf, "max-slack recovery upload");
  }

  explicit RecoveryStep(const presolve::AntipodalComponentPrimalRecoveryStep& a) : antipodal(true) {
    nanti = static_cast<int32_t>(a.plus_cols.size());
    up(a_plus, a.plus_cols, "antipodal recovery upload");
    up(a_minus, a.minus_cols, "antipodal recovery upload");
    up(a_rplus, a.root_plus_cols, "antipodal recovery upload");
    up(a_rminus, a.root_minus_cols, "antipodal recovery upload");
    up(a_plower, a.plus_lower, "antipodal recovery upload");
    up(a_mlower, a.minus_lower, "antipodal recovery upload");
    a_pcache.alloc(nanti);
    a_mcache.alloc(nanti);
  }

  // restore_structural_primal_step / restore_antipodal_component
  // (gpu_postsolve.cu:1150-1212): splits, outer pairs, linked slacks, then
  // max slacks.
  void enqueue(double* x, const double* original_l, hipStream_t s) const {
    if (antipodal) {
      check(gpk_restore_antipodal(x, a_plus, a_minus, a_rplus, a_rminus, a_plower, a_mlower, a_pcache, a_mcache, nanti,
                                  s),
            "antipodal recovery");
      return;
    }
    check(gpk_restore_splits(x, split_t, split_e, split_rho, nsplit, s), "split recovery");
    check(gpk_restore_outer_pairs(x, outer_b, outer_f, nouter, s), "outer-pair recovery");
    check(gpk_restore_linked_slacks(x, linked_s, linked_t, linked_f, nlinked, s), "linked-slack recovery");
    check(gpk_restore_max_slacks(x, max_s, max_start, max_t, max_f, nmax, original_l, s), "max-slack recovery");
  }
};

}  // namespace

void PostsolvePlan::validate(const PresolveRecordGpu& record, const LPInfoGpu* original) {
  validate_recovery_timeline(record);
  if (static_cast<int32_t>(record.col_red2org.size()) != record.n1 ||
      static_cast<int32_t>(record.row_red2org.size()) != record.m1)
    throw std::runtime_error("postsolve_gpu: record maps do not match m1/n1");
  if (record.fixed_idx.size() != record.fixed_val.size())
    throw std::runtime_error("postsolve_gpu: fixed_idx and fixed_val differ in length");
  if (original != nullptr && (original->A.rows != record.m0 || original->A.cols != record.n0))
    throw std::runtime_error("postsolve_gpu: original model does not match the record's dimensions");
}

struct PostsolvePlan::Impl {
  int32_t n0 = 0, m0 = 0, n1 = 0, m1 = 0, nf = 0;
  bool project = false;
  gp_orig_dev orig{};
  std::unique_ptr<presolve::DeviceTape> tape;
  DBuf<int32_t> col_map, row_map, fidx;
  DBuf<double> fval;
  DBuf<uint8_t> zret, prot;
  DBuf<int32_t> long_cols, long_count;
  presolve::Schedule dual;
  DBuf<int32_t> dual_buf;
  // The primal pass in replay order: a schedule over a tape range, then the
  // recovery step at that range's start (none after the last range).
  struct Stage {
    presolve::Schedule sched;
    DBuf<int32_t> buf;
    std::unique_ptr<RecoveryStep> step;
  };
  std::vector<Stage> primal;
  const double* original_l = nullptr;
  int64_t records = 0, tape_bytes = 0;
};

PostsolvePlan::PostsolvePlan(const PresolveRecordGpu& record, const LPInfoGpu* original, const double* original_l)
    : d_(new Impl) {
  AbandonOnThrow abandon_guard;
  validate(record, original);
  if (original != nullptr && original->A.rowPtr != nullptr && original->AT.rowPtr != nullptr &&
      original->AT.rows == original->A.cols && original->AT.nnz == original->A.nnz)
    check_lp_arrays(*original, "postsolve original LP");
  Impl& d = *d_;
  auto& c = ctx();
  hipStream_t s = c.stream();
  d.n0 = record.n0;
  d.m0 = record.m0;
  d.n1 = record.n1;
  d.m1 = record.m1;
  {
    const presolve::PostsolveTape& t = record.tape;
    d.records = static_cast<int64_t>(t.types.size());
    d.tape_bytes = static_cast<int64_t>(sizeof(int32_t) * (t.types.size() + t.index_starts.size() +
                                                            t.value_starts.size() + t.indices.size()) +
                                        t.dual_modes.size() + sizeof(double) * t.vals.size());
  }
  // GP_PRESOLVE_PROFILE=1: host time of the preparation steps on stderr.
  const bool prof = std::getenv("GP_PRESOLVE_PROFILE") != nullptr;
  auto tick = std::chrono::steady_clock::now();
  auto lap = [&](const char* what) {
    if (!prof) return;
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "  postsolve %-28s %8.3f ms\n", what, 1e3 * std::chrono::duration<double>(now - tick).count());
    tick = now;
  };
  d.zret.alloc(d.n0);
  // The record's maps and tape go up together (copies only in this group).
  c.begin_copy_group();
  presolve::upload_map(d.col_map, record.col_red2org, d.n0, "col_red2org upload");
  presolve::upload_map(d.row_map, record.row_red2org, d.m0, "row_red2org upload");
  presolve::upload_map(d.fidx, record.fixed_idx, d.n0, "fixed_idx upload");
  d.nf = static_cast<int32_t>(record.fixed_val.size());
  d.fval.alloc(d.nf);
  c.to_device(d.fval.get(), record.fixed_val.data(), d.nf, "fixed_val upload");
  d.tape = std::make_unique<presolve::DeviceTape>(record.tape);
  c.end_copy_group("record upload");
  d.orig = presolve::orig_view(original);
  const presolve::TapeScanner scanner(record.tape, d.n0, d.m0);
  const bool has_orig = original != nullptr;
  const bool has_y = d.m0 > 0;
  presolve::OrigPattern pattern;
  if (has_orig) {
    std::vector<int32_t> rows, cols;
    scanner.pattern_lines(rows, cols);
    pattern.a_rows = presolve::gather_lines(original->A.rowPtr, original->A.colVal, d.m0, rows);
    pattern.at_cols = presolve::gather_lines(original->AT.rowPtr, original->AT.colVal, d.n0, cols);
  }
  lap("tape upload, scan, lines");
  const int32_t count = d.tape->view.count;
  d.original_l = original_l != nullptr ? original_l : (original != nullptr ? original->l : nullptr);
  // postsolve_gpu (gpu_postsolve.cu:1691-1713): the primal pass replays the
  // tape from its end down to each recovery checkpoint, in reverse
  // chronological order, and applies that checkpoint's recovery step there.
  int32_t range_end = count;
  for (auto it = record.primal_recovery_timeline.rbegin(); it != record.primal_recovery_timeline.rend(); ++it) {
    Impl::Stage st;
    st.sched = presolve::level_schedule(scanner, it->tape_position, range_end, d.n0, d.m0, true, has_orig, has_y,
                                        &pattern);
    st.buf = presolve::upload_schedule(st.sched);
    if (it->kind == presolve::PrimalRecoveryKind::StructuralL1)
      st.step = std::make_unique<RecoveryStep>(record.structural_primal_recoveries[it->payload_index]);
    else
      st.step = std::make_unique<RecoveryStep>(record.antipodal_component_recoveries[it->payload_index]);
    d.primal.push_back(std::move(st));
    range_end = it->tape_position;
  }
  {
    Impl::Stage st;
    st.sched = presolve::level_schedule(scanner, 0, range_end, d.n0, d.m0, true, has_orig, has_y, &pattern);
    st.buf = presolve::upload_schedule(st.sched);
    d.primal.push_back(std::move(st));
  }
  lap("primal schedule");
  d.dual = presolve::level_schedule(scanner, 0, count, d.n0, d.m0, false, has_orig, has_y, &pattern);
  d.dual_buf = presolve::upload_schedule(d.dual);
  lap("dual schedule");
  // project_column_duals_from_original (gpu_postsolve.cu:1370-1435): the
  // columns the tape recovers duals for keep them.
  d.project = original != nullptr && has_y && original->AT.rows > 0 && original->c && original->l && original->u;
  if (d.project) {
    d.prot.alloc(d.n0);
    CopyBatch().zero(d.prot.get(), d.n0).run("protected columns clear");
    check(gpk_protect_columns(d.tape->view, d.n0, d.prot, s), "protected columns");
    d.long_cols.alloc(d.n0);
    d.long_count.alloc(1);
  }
}

PostsolvePlan::~PostsolvePlan() = default;
int32_t PostsolvePlan::n0() const { return d_->n0; }
int32_t PostsolvePlan::m0() const { return d_->m0; }
int32_t PostsolvePlan::n1() const { return d_->n1; }
int32_t PostsolvePlan::m1() const { return d_->m1; }

void PostsolvePlan::enqueue(const double* x_red, const double* y_red, const double* z_red, double* x_org,
                            double* y_org, double* z_org) {
  Impl& d = *d_;
  hipStream_t s = ctx().stream();
  double* y = d.m0 > 0 ? y_org : nullptr;
  CopyBatch batch;
  batch.zero(x_org, d.n0).zero(z_org, d.n0).zero(d.zret.get(), d.n0).zero(y_org, d.m0);
  if (d.project) batch.zero(d.long_count.get(), 1);
  batch.run("postsolve output clear");
  if (x_red) check(gpk_scatter(x_red, d.col_map, d.n1, x_org, s), "scatter x");
  if (y_red && d.m0 > 0) check(gpk_scatter(y_red, d.row_map, d.m1, y_org, s), "scatter y");
  if (z_red) check(gpk_scatter(z_red, d.col_map, d.n1, z_org, s), "scatter z");
  check(gpk_mark(d.col_map, d.n1, d.zret, s), "mark retrieved");
  check(gpk_restore_fixed(d.fidx, d.fval, d.nf, x_org, s), "restore fixed");
  for (const Impl::Stage& st : d.primal) {
    presolve::enqueue_schedule(st.sched, st.buf, d.tape->view, d.orig, x_org, y, z_org, d.zret, true, kPostsolveTol);
    if (st.step) st.step->enqueue(x_org, d.original_l, s);
  }
  presolve::enqueue_schedule(d.dual, d.dual_buf, d.tape->view, d.orig, x_org, y, z_org, d.zret, false,
                             kPostsolveTol);
  if (d.project)
    check(gpk_project_duals(d.orig, d.prot, x_org, y, z_org, d.n0, kPostsolveTol, d.long_count, d.long_cols, s),
          "project duals");
}

PostsolvePlan::Shape PostsolvePlan::shape() const {
  Shape out;
  out.records = d_->records;
  out.tape_bytes = d_->tape_bytes;
  for (const auto& st : d_->primal)
    out.primal_levels += static_cast<int32_t>(st.sched.level_start.empty() ? 0 : st.sched.level_start.size() - 1);
  out.dual_levels = static_cast<int32_t>(d_->dual.level_start.empty() ? 0 : d_->dual.level_start.size() - 1);
  return out;
}

}  // namespace gpu_presolver::amd

namespace gpu_presolver::presolve {

// gpu_postsolve.cu:1618-1628: the steps in reverse order on x_org, then a
// device synchronization.
void postsolve_restore_structural_primal_gpu(double* x_org, const std::vector<StructuralL1PrimalRecoveryStep>& recoveries,
                                             const double* original_l) {
  amd::AbandonOnThrow abandon_guard;
  if (!recoveries.empty()) {
    check(hipDeviceSynchronize(), "structural primal recovery entry synchronize");
    for (auto it = recoveries.rbegin(); it != recoveries.rend(); ++it) {
      amd::RecoveryStep step(*it);
      step.enqueue(x_org, original_l, ctx().stream());
      ctx().sync("structural primal recovery");
    }
  }
  check(hipDeviceSynchronize(), "postsolve_restore_structural_primal_gpu synchronize");
}

GpuPostsolveResult postsolve_gpu(double* x_red, double* y_red, double* z_red, const PresolveRecordGpu& record,
                                 const double* original_l) {
  return postsolve_gpu(x_red, y_red, z_red, record, nullptr, original_l);
}

GpuPostsolveResult postsolve_gpu(double* x_red, double* y_red, double* z_red, const PresolveRecordGpu& record,
                                 const LPInfoGpu* original_model, const double* original_l) {
  amd::AbandonOnThrow abandon_guard;
  amd::PostsolvePlan::validate(record, original_model);
  if (amd::job_mode().device_sync_on_entry) check(hipDeviceSynchronize(), "postsolve entry synchronize");
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
    amd::PostsolvePlan plan(record, original_model, original_l);
    plan.enqueue(x_red, y_red, z_red, res.x_org, res.y_org, res.z_org);
    c.sync("postsolve_gpu synchronize");
  } catch (...) {
    (void)hipStreamSynchronize(s);
    amd::dev_free(res.x_org, s);
    amd::dev_free(res.y_org, s);
    amd::dev_free(res.z_org, s);
    throw;
  }
  return res;
}

}  // namespace gpu_presolver::presolve
