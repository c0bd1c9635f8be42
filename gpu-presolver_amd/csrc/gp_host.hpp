// Host-side infrastructure of libgpu_presolver_amd: HIP error handling, the
// per-thread execution context (stream, flag words, pinned staging) and
// stream-ordered device buffers.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstring>
#include <exception>
#include <functional>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gp_kernels.h"
#include "gp_presolve_capi.h"
#include "gpu_presolver/presolve/presolve_structs.hpp"

namespace gpu_presolver::amd {

// A failed HIP runtime call: std::runtime_error("<context>: <hip error
// string>"), the reference's error convention (presolve_helpers.hpp:13-19).
class HipError : public std::runtime_error {
 public:
  HipError(const char* context, hipError_t status)
      : std::runtime_error(std::string(context) + ": " + hipGetErrorString(status)), status_(status) {}
  hipError_t status() const { return status_; }

 private:
  hipError_t status_;
};

// An input that breaks the LP contract (bad CSR arrays, mismatched sizes):
// a std::runtime_error like every other failure of the reference's API; the
// C-ABI reports it as GP_ERR_INVALID_ARGUMENT.
class InputError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

// The LP has a block structure on which one of the structure-specific rules
// would act; those rules are not implemented, so the presolve stops instead
// of returning a record that differs from the reference's.  The C-ABI
// reports it as GP_ERR_UNSUPPORTED.
class UnsupportedStructure : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

inline void check(hipError_t status, const char* context) {
  if (status == hipSuccess) return;
  throw HipError(context, status);
}

// Execution context: one stream, a small block of device flag words read
// back through pinned memory, and a device-memory pool that keeps freed
// blocks cached so per-rule scratch allocation stays cheap.
class Ctx {
 public:
  static constexpr int kFlags = 16;
  static constexpr int kFlagSlots = 256;
  explicit Ctx(hipStream_t stream = nullptr, bool own_stream = true);
  ~Ctx();
  Ctx(const Ctx&) = delete;
  Ctx& operator=(const Ctx&) = delete;

  hipStream_t stream() const { return stream_; }
  // Workspace of one single-pass (look-back) launch over `tiles` tiles: the
  // persistent status array (grown on demand), a fresh zeroed ticket word and
  // a new epoch.  The error flag is checked by sync().
  gpk_lookback lookback(int64_t tiles);
  // Returns kFlags zeroed device flag words.
  int32_t* flags();
  // Copies the first `count` words of the flags last returned by flags() to
  // the host (waits for the stream).  The words stay valid until the next
  // read_flags, scalar_to_host or mail call on this context.
  const int32_t* read_flags(int count = kFlags);
  // Waits for everything enqueued on the stream so far.
  void sync(const char* context = "stream synchronize");
  // Enqueues one small kernel that copies `words` (<= kMailWords) 32-bit
  // words from device memory into host-mapped memory and then stores a new
  // sequence number there; the host polls that word.  Returns the words once
  // the whole stream up to that kernel has finished.  This avoids a copy
  // command plus hipStreamSynchronize per decision, which is what limits
  // many presolve jobs sharing one GPU.
  const uint32_t* mail(const void* dev, int words, const char* context);
  static constexpr int kMailWords = 16;
  // Number of host synchronizations so far (each read_flags / to_host /
  // sync counts one); the presolve step profiler reports it per step.
  int64_t sync_count() const { return syncs_; }
  // Seconds the host has spent waiting for the stream (mail polls and
  // stream synchronizations).
  double wait_seconds() const { return wait_s_; }

  // Enqueues a copy of `bytes` of device memory into host memory at `dst`:
  // a kernel writes it into the context's host-mapped staging area and the
  // next wait (sync, read_flags, scalar_to_host, mail) moves it to `dst`.
  // `dst` must stay valid until then.
  void download(void* dst, const void* dev, size_t bytes, const char* context);
  template <class T>
  std::vector<T> to_host(const T* dev, int64_t count, const char* context) {
    std::vector<T> out(static_cast<size_t>(count > 0 ? count : 0));
    if (count > 0) {
      download(out.data(), dev, sizeof(T) * static_cast<size_t>(count), context);
      sync(context);
    }
    return out;
  }
  // Enqueues a copy into `out` (resized to count); the data is valid after
  // the next sync(), so several copies can share one synchronization.
  template <class T>
  void to_host_async(std::vector<T>& out, const T* dev, int64_t count, const char* context) {
    out.resize(static_cast<size_t>(count > 0 ? count : 0));
    if (count > 0) download(out.data(), dev, sizeof(T) * static_cast<size_t>(count), context);
  }
  template <class T>
  T scalar_to_host(const T* dev, const char* context) {
    static_assert(sizeof(T) % 4 == 0 && sizeof(T) <= 4 * kMailWords, "scalar of whole 32-bit words");
    T v;
    std::memcpy(&v, mail(dev, static_cast<int>(sizeof(T) / 4), context), sizeof(T));
    return v;
  }
  // Copies host memory into the context's host-mapped staging area at once
  // (the source may be released on return) and enqueues the kernel that moves
  // it to device memory at `dev`.
  void upload(void* dev, const void* host, size_t bytes, const char* context);
  // Copies host memory into the staging area and returns its device address,
  // for a kernel to read (and rewrite) in place over the bus; with
  // `dst_after`, the area's first `dst_bytes` bytes, as the kernels left
  // them, are copied there at the next wait.  The address is valid until
  // that wait.
  void* stage(const void* host, size_t bytes, void* dst_after, size_t dst_bytes, const char* context);
  template <class T>
  void to_device(T* dev, const T* host, int64_t count, const char* context) {
    if (count > 0) upload(dev, host, sizeof(T) * static_cast<size_t>(count), context);
  }
  // Staged copies (download, to_host_async, upload, to_device) issued
  // between begin_copy_group() and end_copy_group() go out together, one
  // k_copy_many launch per GPK_COPY_MAX of them, instead of one launch each.
  // Nothing else may be launched on the stream inside a group (the copies
  // are enqueued at its end or when a group fills); a wait (sync,
  // read_flags, scalar_to_host, mail) also ends the group, its copies first.
  void begin_copy_group() { grouping_ = true; }
  void end_copy_group(const char* context);
  // Host work run in the order given once the downloads of the next wait
  // have landed, before that wait returns.  A callback may copy from the
  // (then idle) stream synchronously but must not wait through this context.
  void after_wait(std::function<void()> fn) { after_.push_back(std::move(fn)); }
  bool has_after_wait() const { return !after_.empty(); }
  // Set while the presolve scheduler runs: rules then leave their objective
  // shifts to the next wait (add_ordered_masked_sum).  The scheduler reads
  // a plan's shift only after a wait.
  bool defer_objective_sums = false;
  // Drops every pending download without copying it (its destination may
  // already be gone), waits for the stream so the staging area is free, and
  // clears the spin-limit error word.  Called when an exception leaves a
  // library call, so the next wait on this context does not write into
  // released host memory.
  void abandon() noexcept;

 private:
  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  int32_t* d_flags_ = nullptr;
  int32_t* cur_ = nullptr;
  int slot_ = 0;
  static constexpr int kTickets = 1024;
  int32_t* d_tickets_ = nullptr;
  int ticket_ = 0;
  uint64_t* d_status_ = nullptr;
  int64_t status_cap_ = 0;
  uint32_t epoch_ = 0;
  int64_t syncs_ = 0;
  double wait_s_ = 0.0;
  uint32_t* h_mail_ = nullptr;  // host-mapped: [0] sequence, [4, 4 + kMailWords) payload
  uint32_t* d_mail_ = nullptr;
  uint32_t mail_seq_ = 0;
  struct Download {
    void* dst;
    size_t offset, bytes;
  };
  char* h_stage_ = nullptr;  // host-mapped staging area of download() and upload()
  char* d_stage_ = nullptr;
  size_t stage_cap_ = 0, stage_used_ = 0;
  std::vector<Download> downloads_;
  bool grouping_ = false;
  std::vector<gpk_copy_seg> pending_;  // copies of the open group
  std::vector<std::function<void()>> after_;
  void enqueue_copy(const gpk_copy_seg& seg, const char* context);
  void flush_pending(const char* context);
  void check_error_flag(const char* context);
  void finish_downloads();
  // Offset of `bytes` of free staging space (waits for the stream and grows
  // the area when it is full); released at the next wait.
  size_t stage_reserve(size_t bytes, const char* context);
  int32_t* h_error_ = nullptr;  // host-mapped
  int32_t* d_error_ = nullptr;
};

// Context of the calling thread.  ScopedCtx installs another one (the
// serving and batched drivers run each job on its own stream).
Ctx& ctx();
class ScopedCtx {
 public:
  explicit ScopedCtx(Ctx& c);
  ~ScopedCtx();

 private:
  Ctx* prev_;
};

// Calls abandon() on the calling thread's context, if it has one.
void abandon_pending_copies() noexcept;

// Placed at the top of every public entry point: when an exception leaves
// the scope, the context's pending downloads (whose destinations are host
// vectors being destroyed by the unwinding) are dropped.
class AbandonOnThrow {
 public:
  AbandonOnThrow() : depth_(std::uncaught_exceptions()) {}
  ~AbandonOnThrow() {
    if (std::uncaught_exceptions() > depth_) abandon_pending_copies();
  }
  AbandonOnThrow(const AbandonOnThrow&) = delete;
  AbandonOnThrow& operator=(const AbandonOnThrow&) = delete;

 private:
  int depth_;
};

// Caching stream-ordered allocation (host_common.cpp).  dev_free returns the
// library's own blocks to the cache of their stream and hipFree()s memory the
// caller allocated with hipMalloc (a plan's new_A may be either).
void* dev_alloc(size_t bytes, hipStream_t stream, const char* context);
void dev_free(void* p, hipStream_t stream) noexcept;
// Returns the cached free blocks of `stream` to the device pool.
void release_cached_blocks(hipStream_t stream) noexcept;

// Stream-ordered device buffer (hipMallocAsync / hipFreeAsync).
template <class T>
class DBuf {
 public:
  DBuf() = default;
  explicit DBuf(int64_t count, const char* context = "device buffer") { alloc(count, context); }
  ~DBuf() { reset(); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p_(o.p_), n_(o.n_), s_(o.s_) { o.p_ = nullptr; o.n_ = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      reset();
      p_ = o.p_; n_ = o.n_; s_ = o.s_;
      o.p_ = nullptr; o.n_ = 0;
    }
    return *this;
  }
  void alloc(int64_t count, const char* context = "device buffer") {
    reset();
    s_ = ctx().stream();
    n_ = count;
    const size_t bytes = sizeof(T) * static_cast<size_t>(count > 0 ? count : 1);
    p_ = static_cast<T*>(dev_alloc(bytes, s_, context));
  }
  void reset() noexcept {
    if (p_) dev_free(p_, s_);
    p_ = nullptr;
    n_ = 0;
  }
  // Gives up ownership (the pointer is then released with dev_free / hipFree).
  T* release() noexcept {
    T* p = p_;
    p_ = nullptr;
    n_ = 0;
    return p;
  }
  T* get() const { return p_; }
  operator T*() const { return p_; }
  int64_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  int64_t n_ = 0;
  hipStream_t s_ = nullptr;
};

// Device copies and zero fills collected and issued together: one launch
// per GPK_COPY_MAX ranges on the context's stream (in place of one copy or
// fill command each).
class CopyBatch {
 public:
  template <class T>
  CopyBatch& copy(T* dst, const T* src, int64_t n) {
    if (n > 0) segs_.push_back(gpk_copy_seg{dst, src, static_cast<int64_t>(sizeof(T)) * n});
    return *this;
  }
  template <class T>
  CopyBatch& zero(T* dst, int64_t n) {
    if (n > 0) segs_.push_back(gpk_copy_seg{dst, nullptr, static_cast<int64_t>(sizeof(T)) * n});
    return *this;
  }
  void run(const char* context) {
    for (size_t k = 0; k < segs_.size(); k += GPK_COPY_MAX) {
      const size_t cnt = segs_.size() - k < GPK_COPY_MAX ? segs_.size() - k : GPK_COPY_MAX;
      check(gpk_copy_many(segs_.data() + k, static_cast<int32_t>(cnt), ctx().stream()), context);
    }
    segs_.clear();
  }

 private:
  std::vector<gpk_copy_seg> segs_;
};

// ---- prepared postsolve ------------------------------------------------------
// Everything a postsolve of one record needs that does not depend on the
// reduced point: the validated, uploaded tape and record maps, the two replay
// schedules (built on the host and uploaded), the protected-column mask and
// the work buffers.  enqueue() only launches kernels on the context's stream
// (no allocation, copy or wait), so it can be captured into a graph.
class PostsolvePlan {
 public:
  // Throws std::runtime_error when the record or the original LP cannot be
  // postsolved (the checks of postsolve_gpu).
  static void validate(const presolve::PresolveRecordGpu& record, const presolve::LPInfoGpu* original);
  // original_l: the original lower bounds the max-slack recovery reads
  // (null: original->l, or none).
  PostsolvePlan(const presolve::PresolveRecordGpu& record, const presolve::LPInfoGpu* original,
                const double* original_l = nullptr);
  ~PostsolvePlan();
  PostsolvePlan(const PostsolvePlan&) = delete;
  PostsolvePlan& operator=(const PostsolvePlan&) = delete;
  // x_red, y_red, z_red may be null (as in postsolve_gpu); outputs are device
  // arrays of n0, m0 and n0 doubles.
  void enqueue(const double* x_red, const double* y_red, const double* z_red, double* x_org, double* y_org,
               double* z_org);
  int32_t n0() const;
  int32_t m0() const;
  int32_t n1() const;
  int32_t m1() const;
  // Tape records, dependency levels of the primal pass (over all its ranges)
  // and of the dual pass, and the tape's bytes on the device.
  struct Shape {
    int64_t records = 0;
    int32_t primal_levels = 0, dual_levels = 0;
    int64_t tape_bytes = 0;
  };
  Shape shape() const;

 private:
  struct Impl;
  std::unique_ptr<Impl> d_;
};

// ---- views -----------------------------------------------------------------
inline gpk_csr csr_view(const presolve::DeviceCsrMatrix& a) {
  return gpk_csr{a.rows, a.cols, a.nnz, a.rowPtr, a.colVal, a.nzVal};
}
inline gpk_plan plan_view(const presolve::PresolvePlanGpu& p) {
  return gpk_plan{p.keep_row_mask, p.keep_col_mask, p.new_c, p.new_l, p.new_u, p.new_AL, p.new_AU};
}

// Options of the calling thread's presolve and postsolve calls.  The public
// entry points synchronize the whole device on entry (their inputs may have
// been written on the legacy stream) and hand out hipMalloc'ed outputs that
// the caller releases with hipFree.  A job of the serving runtime has its
// inputs ready, runs on its own stream and returns its outputs to the pool
// (dev_free), so it does neither: a device-wide synchronization or a hipFree
// would stall every other job in flight.
struct JobMode {
  bool device_sync_on_entry = true;
  bool pooled_outputs = false;
};
JobMode& job_mode();

// Serving runtime (host_serve.cpp).
class Server;
Server* server_create(const presolve::PresolveParams& prm, const ::gp_serve_config& cfg);
void server_destroy(Server* s);
int64_t server_submit(Server* s, const presolve::LPInfoGpu& lp, uint64_t tag, double t_arrival);
int64_t server_submit_spec(Server* s, const gpk_lpgen_spec& spec, uint64_t tag, double t_arrival);
// Seeded LP generated on the device into blocks of the calling thread's pool
// (gp_generate_lp); released with free_generated_lp on the same thread.
void check_spec(const gpk_lpgen_spec& spec);
presolve::LPInfoGpu generate_lp(const gpk_lpgen_spec& spec);
void free_generated_lp(presolve::LPInfoGpu& lp) noexcept;
// run_job on the LP of spec, generated first; *t_ready gets the host's
// steady-clock seconds once the LP is resident.
gp_job_result run_generated_job(const gpk_lpgen_spec& spec, const presolve::PresolveParams& prm, bool postsolve,
                                uint64_t seed, uint64_t tag, double* t_ready);
int server_wait(Server* s, ::gp_job_result* out, int max, double timeout_s);
double server_now(const Server* s);
uint64_t checksum_xyz(const double* x, int32_t n, const double* y, int32_t m, const double* z);
// One presolve (+ seeded point + postsolve + checksum) job on the calling
// thread's context, in the thread's current JobMode.
gp_job_result run_job(const presolve::LPInfoGpu& lp, const presolve::PresolveParams& prm, bool postsolve,
                      uint64_t seed, uint64_t tag);

// Working-LP nnz from which each fast phase opens with the batched singleton
// closure (process-wide; 10,000,000 as in the reference).
void set_batched_closure_min_nnz(int32_t nnz);
int32_t batched_closure_min_nnz();
// Testing hook: the calling thread's count-th record update from now throws
// std::runtime_error right after it has queued its host copies (0 disarms).
void fail_after_record_updates(int32_t count);
void record_update_checkpoint();
// Testing hook: on the calling thread, every phase that takes over the last
// phase's plan first compares it (and its statistics) with a freshly built
// pair and throws std::logic_error on any difference.
void check_plan_reuse(bool on);
bool plan_reuse_checked();
// Phases of the calling thread that took over their predecessor's plan.
void count_plan_reuse();
int64_t plan_reuses();

// ---- shared host helpers (host_common.cpp) ---------------------------------
// Row activity (gpu_presolve_kernels.cu:18-111) on `stream`: the bounds are
// first interleaved into a pool block (one 16-byte gather per entry), then
// k_row_activity_lu runs; enqueue only.
void row_activity(const int32_t* ptr, const int32_t* idx, const double* val, int32_t m, int32_t n, const double* l,
                  const double* u, double ztol, double* min_fin, double* max_fin, int32_t* neg_inf, int32_t* pos_inf,
                  hipStream_t stream);
// Exclusive scan of a uint8 mask into out[0..n]; returns the count.
int32_t scan_mask_count(const uint8_t* mask, int32_t n, int32_t* out);
// The same scans without reading the total back (out[n] holds it).
void scan_mask_launch(const uint8_t* mask, int32_t n, int32_t* out);
void scan_counts_launch(const int32_t* in, int32_t n, int32_t* out);
// Exclusive scan of int32 counts into out[0..n]; returns the total.
int32_t scan_counts(const int32_t* in, int32_t n, int32_t* out);
// Sum, in ascending index order, of vals[i] over mask[i] != 0 (computed on
// the host from the packed selection, so the rounding order is fixed).
double ordered_masked_sum(const uint8_t* mask, const double* vals, int32_t n);
// *target += ordered_masked_sum(mask, vals, n).  With the context's
// defer_objective_sums set, the terms are packed and downloaded without a
// wait and the addition runs at the next wait (Ctx::after_wait), in the
// order of the calls; the mask and values may be released once this returns.
void add_ordered_masked_sum(double* target, const uint8_t* mask, const double* vals, int32_t n);
// Compacts A by row/column keep maps; the result owns fresh device arrays.
// known_nnz >= 0 is the kept count when the caller knows it (the transpose
// of a matrix compacted with the same maps): the read-back is skipped.
// col_rank: the column map's rank words when the caller has them (unused
// when n_new == A.cols: no column is dropped and col_map is the identity).
presolve::DeviceCsrMatrix compact_csr(const presolve::DeviceCsrMatrix& A, const int32_t* row_map,
                                      const int32_t* col_map, int32_t m_new, int32_t n_new, int32_t known_nnz = -1,
                                      const gpk_rank_word* col_rank = nullptr);
// The same compaction in two halves, for a caller that reads the new sizes
// in its own wait: compact_start launches the map pass (and, past 256 tiles,
// the tile-offset scan) while the sizes are still unknown (every kept column
// goes through col_rank, which must be given); the map pass adds the kept
// entry count to the device word `kept`, which the caller zeroes.
// compact_finish then writes the compacted matrix.  The row-pointer array is
// sized for A.rows + 1.
struct CompactStart {
  const presolve::DeviceCsrMatrix* A = nullptr;
  const int32_t* row_map = nullptr;
  DBuf<int32_t> ptr, jn, tile_off;
  bool few_tiles = false;  // tile_off holds raw counts (no scan launched)
};
CompactStart compact_start(const presolve::DeviceCsrMatrix& A, const int32_t* row_map, const gpk_rank_word* col_rank,
                           int32_t* kept);
presolve::DeviceCsrMatrix compact_finish(CompactStart& st, int32_t m_new, int32_t n_new, int32_t nnz_new);
// The two halves for A and a transpose of it with sorted rows, compacted by
// the same maps (AT's rows by col_map, its columns by the row rank words):
// compact_start_pair launches one map pass over both matrices' tiles and, past
// 256 tiles on either side, one scan of both sides' tile counts;
// compact_finish_pair writes both in one launch.  `kept` receives A's kept
// count as in compact_start.
struct CompactPairStart {
  const presolve::DeviceCsrMatrix* A = nullptr;
  const presolve::DeviceCsrMatrix* AT = nullptr;
  const int32_t* row_map = nullptr;
  const int32_t* col_map = nullptr;
  DBuf<int32_t> ptr, t_ptr, jn, t_jn, tile_off;
  bool from_counts = false;  // tile_off holds raw counts (no scan launched)
};
CompactPairStart compact_start_pair(const presolve::DeviceCsrMatrix& A, const int32_t* row_map,
                                    const gpk_rank_word* col_rank, const presolve::DeviceCsrMatrix& AT,
                                    const int32_t* col_map, const gpk_rank_word* row_rank, int32_t* kept);
// {compacted A, compacted AT}; both hold nnz_new entries.
std::pair<presolve::DeviceCsrMatrix, presolve::DeviceCsrMatrix> compact_finish_pair(CompactPairStart& st,
                                                                                     int32_t m_new, int32_t n_new,
                                                                                     int32_t nnz_new);
// Stable transpose (rows of the result sorted); fresh device arrays.
presolve::DeviceCsrMatrix transpose_csr(const presolve::DeviceCsrMatrix& A);
void free_csr(presolve::DeviceCsrMatrix& A) noexcept;
// Sorts the rows of A in place by column when some row is unsorted.
void sort_rows_if_needed(const presolve::DeviceCsrMatrix& A);
// Checks A and AT of an LP on the device (row pointers, index range, no
// column repeated within a row; one wait) and throws InputError naming the
// first problem.  Unsorted rows are accepted.
void check_lp_arrays(const presolve::LPInfoGpu& lp, const char* who);
// Fail-closed screens of the structure-specific rules (host_screens.cpp):
// each throws UnsupportedStructure when its rule would act on the LP.
void screen_bootstrap_structures(const presolve::LPInfoGpu& lp, const presolve::PresolveParams& prm);

// ---- on-device LP solver (host_pdhg.cpp) -----------------------------------
// Restarted PDHG on the Ruiz-scaled LP, with every vector resident on the
// device; the host only reads the KKT sums once per check period.
struct PdhgOptions {
  int32_t max_iters = 200000;
  double max_time = 3600.0;
  double tol = 1e-4;          // relative KKT tolerance
  int32_t check_every = 64;   // iterations per check period (one graph launch)
  int32_t ruiz_iters = 10;
  int32_t power_iters = 40;
  int32_t use_graph = 1;
  int32_t verbose = 0;
  int32_t method = 0;  // 0 restarted averaged PDHG, 1 restarted reflected Halpern PDHG
};
struct PdhgReport {
  int32_t status = 0;  // 0 optimal, 1 iteration limit, 2 time limit
  int32_t iterations = 0, restarts = 0, checks = 0;
  double primal_obj = 0.0, dual_obj = 0.0;
  double primal_res = 0.0, dual_res = 0.0, gap = 0.0;  // relative
  double elapsed_seconds = 0.0, norm_estimate = 0.0, primal_weight = 0.0;
  int64_t graph_launches = 0;
};
// x, z: device arrays of lp.A.cols doubles, y of lp.A.rows; on return they
// hold the best iterate (unscaled), z the reduced costs projected on the
// cone of the column bounds.
PdhgReport pdhg_solve(const presolve::LPInfoGpu& lp, const PdhgOptions& opt, double* x, double* y, double* z);

// ---- presolve + solve + postsolve (host_pipeline.cpp) ----------------------
}  // namespace gpu_presolver::amd
namespace gpu_presolver::presolve {
struct GpuPresolveSummary;
}  // namespace gpu_presolver::presolve
namespace gpu_presolver::amd {
struct PipelineReport {
  int32_t reduced_rows = 0, reduced_cols = 0, reduced_nnz = 0, iterations = 0;
  double obj_constant_delta = 0.0;
  bool has_infeasible = false, has_unbounded = false;
  PdhgReport solve;  // status -1: not solved (presolve found the LP infeasible or unbounded)
  double presolve_seconds = 0.0, solve_seconds = 0.0, prepare_seconds = 0.0, postsolve_seconds = 0.0,
         total_seconds = 0.0;
  void copy_summary(const presolve::GpuPresolveSummary& s);
};
// Presolves lp, solves the reduced LP in device memory with pdhg_solve while
// a second thread and stream prepare the postsolve, then postsolves into x,
// z (lp.A.cols doubles) and y (lp.A.rows), device arrays.
PipelineReport presolve_solve(const presolve::LPInfoGpu& lp, const presolve::PresolveParams& prm,
                              const PdhgOptions& solve, double* x, double* y, double* z);
void screen_structural_l1(const presolve::LPInfoGpu& lp, const presolve::PresolveParams& prm);
// old_to_new map of a keep mask; returns the kept count.
int32_t build_old_to_new(const uint8_t* mask, int32_t n, DBuf<int32_t>& map);
// FixedCol records and the long-double objective delta
// (fixed_col_tape.cu:107-267).
double fixed_col_tape(presolve::PostsolveTape* tape, const uint8_t* fixed_mask, const uint8_t* excluded,
                      const double* fixed_val, const uint8_t* keep_row, const double* c,
                      const presolve::DeviceCsrMatrix& AT, int32_t n);
void append_record(presolve::PostsolveTape& tape, presolve::PostsolveReductionType type,
                   const int32_t* idx, size_t ni, const double* vals, size_t nv, presolve::PostsolveDualMode mode);
inline void append_record(presolve::PostsolveTape& tape, presolve::PostsolveReductionType type,
                          std::initializer_list<int32_t> idx, std::initializer_list<double> vals,
                          presolve::PostsolveDualMode mode) {
  append_record(tape, type, idx.begin(), idx.size(), vals.begin(), vals.size(), mode);
}

}  // namespace gpu_presolver::amd

namespace gpu_presolver::presolve {
// apply_rule_doubleton_equations that also hands back the transpose (rows
// sorted) of the A it rebuilds, in library blocks the caller frees with
// free_csr; new_at->rowPtr stays null when A is not rebuilt.
void apply_rule_doubleton_equations_with_at(PresolvePlanGpu& plan, const LPInfoGpu& lp, const PresolveParams& prm,
                                            DeviceCsrMatrix* new_at);
// apply_rule_infeasible_fixed_variables, with the empty-column rule's
// detection launched beside it on the same plan and both read in one wait.
// Returns true when the empty-column rule is settled as well (the fixed-
// variable rule fixed nothing, so its detection saw the plan it would have
// seen after it); false when the caller still runs apply_rule_empty_cols.
bool apply_rule_fixed_variables_and_empty_cols(PresolvePlanGpu& plan, const LPInfoGpu& lp,
                                               const PresolveStatsGpu& stats, const PresolveParams& prm);
}  // namespace gpu_presolver::presolve
