// Host-side infrastructure: context, device memory, scans, compaction,
// transpose and the FixedCol tape packer.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gp_host.hpp"

namespace gpu_presolver::amd {

using presolve::DeviceCsrMatrix;
using presolve::PostsolveDualMode;
using presolve::PostsolveReductionType;
using presolve::PostsolveTape;

namespace {

// Keep freed blocks of the device's default pool cached for reuse: rules
// allocate and release their scratch once per call.
void configure_pool_once() {
  static std::once_flag once;
  std::call_once(once, []() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess || pool == nullptr) return;
    uint64_t threshold = std::numeric_limits<uint64_t>::max();
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
  });
}

thread_local Ctx* t_current = nullptr;
thread_local Ctx* t_own = nullptr;  // the thread's own context (created on first use)

}  // namespace

Ctx::Ctx(hipStream_t stream, bool own_stream) {
  configure_pool_once();
  if (stream == nullptr && own_stream) {
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    own_stream_ = true;
  } else {
    stream_ = stream;
    own_stream_ = false;
  }
  check(hipMalloc(&d_flags_, sizeof(int32_t) * kFlags * kFlagSlots), "hipMalloc flags");
  check(hipMemsetAsync(d_flags_, 0, sizeof(int32_t) * kFlags * kFlagSlots, stream_), "hipMemset flags");
  check(hipMalloc(&d_tickets_, sizeof(int32_t) * kTickets), "hipMalloc tickets");
  check(hipMemsetAsync(d_tickets_, 0, sizeof(int32_t) * kTickets, stream_), "hipMemset tickets");
  check(hipHostMalloc(&h_error_, sizeof(int32_t), hipHostMallocMapped), "hipHostMalloc error flag");
  *h_error_ = 0;
  check(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_error_), h_error_, 0), "hipHostGetDevicePointer");
  check(hipHostMalloc(&h_mail_, sizeof(uint32_t) * (4 + kMailWords), hipHostMallocMapped | hipHostMallocCoherent),
        "hipHostMalloc mail");
  std::memset(h_mail_, 0, sizeof(uint32_t) * (4 + kMailWords));
  check(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_mail_), h_mail_, 0), "hipHostGetDevicePointer mail");
}

gpk_lookback Ctx::lookback(int64_t tiles) {
  if (tiles > status_cap_) {
    int64_t cap = 1024;
    while (cap < tiles) cap <<= 1;
    sync("look-back status growth");
    if (d_status_) check(hipFree(d_status_), "hipFree look-back status");
    d_status_ = nullptr;
    check(hipMalloc(&d_status_, sizeof(uint64_t) * static_cast<size_t>(cap)), "hipMalloc look-back status");
    check(hipMemsetAsync(d_status_, 0, sizeof(uint64_t) * static_cast<size_t>(cap), stream_),
          "hipMemset look-back status");
    status_cap_ = cap;
  }
  if (ticket_ == kTickets) {
    check(hipMemsetAsync(d_tickets_, 0, sizeof(int32_t) * kTickets, stream_), "hipMemset tickets");
    ticket_ = 0;
  }
  epoch_ = epoch_ + 1 >= (1u << 30) ? 1u : epoch_ + 1;
  return gpk_lookback{d_status_, d_tickets_ + ticket_++, d_error_, epoch_};
}

Ctx::~Ctx() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  release_cached_blocks(stream_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  (void)hipFree(d_flags_);
  (void)hipFree(d_tickets_);
  if (d_status_) (void)hipFree(d_status_);
  (void)hipHostFree(h_error_);
  (void)hipHostFree(h_mail_);
  if (h_stage_) (void)hipHostFree(h_stage_);
  if (own_stream_) (void)hipStreamDestroy(stream_);
}

// Flag words come from a ring of zeroed slots; the whole ring is cleared
// (stream-ordered, after every earlier use) once per kFlagSlots handouts.
int32_t* Ctx::flags() {
  if (slot_ == kFlagSlots) {
    check(hipMemsetAsync(d_flags_, 0, sizeof(int32_t) * kFlags * kFlagSlots, stream_), "hipMemset flags");
    slot_ = 0;
  }
  cur_ = d_flags_ + static_cast<size_t>(slot_++) * kFlags;
  return cur_;
}

const int32_t* Ctx::read_flags(int count) {
  static_assert(kFlags <= kMailWords, "a flag slot fits the mailbox");
  if (count < 0 || count > kFlags) throw std::invalid_argument("read_flags: count out of range");
  return reinterpret_cast<const int32_t*>(mail(cur_ ? cur_ : d_flags_, count, "read flags"));
}

void Ctx::check_error_flag(const char* context) {
  if (*h_error_ != 0) {
    *h_error_ = 0;
    throw std::runtime_error(std::string(context) + ": a single-pass scan exceeded its spin limit");
  }
}

void Ctx::sync(const char* context) { (void)mail(nullptr, 0, context); }

size_t Ctx::stage_reserve(size_t bytes, const char* context) {
  size_t off = (stage_used_ + 255) & ~static_cast<size_t>(255);
  if (off + bytes > stage_cap_) {
    if (stage_used_ > 0) sync(context);  // staged copies in flight complete (and the area is released)
    off = 0;
    if (bytes > stage_cap_) {
      if (h_stage_) check(hipHostFree(h_stage_), "hipHostFree staging");
      h_stage_ = d_stage_ = nullptr;
      stage_cap_ = 0;
      size_t cap = size_t{1} << 20;
      while (cap < bytes) cap <<= 1;
      check(hipHostMalloc(reinterpret_cast<void**>(&h_stage_), cap, hipHostMallocMapped | hipHostMallocCoherent),
            "hipHostMalloc staging");
      check(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_stage_), h_stage_, 0), "staging device pointer");
      stage_cap_ = cap;
    }
  }
  stage_used_ = off + bytes;
  return off;
}

// Copies larger than this bypass the staging area (a direct copy and a
// stream synchronization), so the pinned area stays modest.
constexpr size_t kStageMaxBytes = size_t{64} << 20;

void Ctx::download(void* dst, const void* dev, size_t bytes, const char* context) {
  if (bytes == 0) return;
  if (bytes > kStageMaxBytes) {
    check(hipMemcpyAsync(dst, dev, bytes, hipMemcpyDeviceToHost, stream_), context);
    check(hipStreamSynchronize(stream_), context);
    return;
  }
  const size_t off = stage_reserve(bytes, context);
  enqueue_copy(gpk_copy_seg{d_stage_ + off, dev, static_cast<int64_t>(bytes)}, context);
  downloads_.push_back(Download{dst, off, bytes});
}

void Ctx::enqueue_copy(const gpk_copy_seg& seg, const char* context) {
  if (!grouping_) {
    check(gpk_copy_many(&seg, 1, stream_), context);
    return;
  }
  pending_.push_back(seg);
  if (pending_.size() == GPK_COPY_MAX) flush_pending(context);
}

void Ctx::flush_pending(const char* context) {
  if (pending_.empty()) return;
  const hipError_t e = gpk_copy_many(pending_.data(), static_cast<int32_t>(pending_.size()), stream_);
  pending_.clear();
  check(e, context);
}

void Ctx::end_copy_group(const char* context) {
  grouping_ = false;
  flush_pending(context);
}

void* Ctx::stage(const void* host, size_t bytes, void* dst_after, size_t dst_bytes, const char* context) {
  if (bytes == 0) return nullptr;
  if (dst_bytes > bytes) throw std::invalid_argument("Ctx::stage: dst_bytes exceeds the staged bytes");
  if (bytes > kStageMaxBytes) throw std::length_error("Ctx::stage: more than 64 MB to stage");
  const size_t off = stage_reserve(bytes, context);
  std::memcpy(h_stage_ + off, host, bytes);
  if (dst_after && dst_bytes > 0) downloads_.push_back(Download{dst_after, off, dst_bytes});
  return d_stage_ + off;
}

void Ctx::upload(void* dev, const void* host, size_t bytes, const char* context) {
  if (bytes == 0) return;
  if (bytes > kStageMaxBytes) {
    check(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, stream_), context);
    check(hipStreamSynchronize(stream_), context);
    return;
  }
  const size_t off = stage_reserve(bytes, context);
  std::memcpy(h_stage_ + off, host, bytes);
  enqueue_copy(gpk_copy_seg{dev, d_stage_ + off, static_cast<int64_t>(bytes)}, context);
}

void Ctx::finish_downloads() {
  for (const Download& d : downloads_) std::memcpy(d.dst, h_stage_ + d.offset, d.bytes);
  downloads_.clear();
  stage_used_ = 0;
  if (after_.empty()) return;
  std::vector<std::function<void()>> run;
  run.swap(after_);
  for (auto& fn : run) fn();
}

const uint32_t* Ctx::mail(const void* dev, int words, const char* context) {
  grouping_ = false;  // a wait closes the group; its copies precede the wait
  flush_pending(context);
  ++syncs_;
  if (++mail_seq_ == 0) mail_seq_ = 1;
  const uint32_t seq = mail_seq_;
  check(gpk_mail(static_cast<const uint32_t*>(dev), words, d_mail_, seq, stream_), context);
  volatile uint32_t* word = h_mail_;
  const auto t0 = std::chrono::steady_clock::now();
  // Spin briefly (most decisions wait a few microseconds), then yield; every
  // few thousand polls the stream is asked for an error so a failed launch
  // does not leave the host waiting.
  for (int64_t polls = 0; word[0] != seq; ++polls) {
    if (polls < 2048) {
      __builtin_ia32_pause();
      continue;
    }
    if ((polls & 4095) == 0) {
      const hipError_t e = hipStreamQuery(stream_);
      if (e != hipSuccess && e != hipErrorNotReady) throw HipError(context, e);
    }
    std::this_thread::yield();
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  wait_s_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  finish_downloads();
  check_error_flag(context);
  return const_cast<const uint32_t*>(h_mail_ + 4);
}

Ctx& ctx() {
  if (t_current) return *t_current;
  if (!t_own) t_own = new Ctx();  // lives for the thread
  return *t_own;
}

void Ctx::abandon() noexcept {
  if (stream_) (void)hipStreamSynchronize(stream_);
  grouping_ = false;
  pending_.clear();
  downloads_.clear();
  after_.clear();
  stage_used_ = 0;
  if (h_error_) *h_error_ = 0;
}

void abandon_pending_copies() noexcept {
  Ctx* c = t_current ? t_current : t_own;
  if (c) c->abandon();
}

JobMode& job_mode() {
  thread_local JobMode mode;
  return mode;
}

ScopedCtx::ScopedCtx(Ctx& c) : prev_(t_current) { t_current = &c; }
ScopedCtx::~ScopedCtx() { t_current = prev_; }

// ---- caching device allocator ----------------------------------------------
// Blocks are rounded up to a power of two (>= 512 B) and, when released, kept
// on a free list of the stream they were allocated on.  Work on one stream
// runs in order, so a block released after its last use can be handed to the
// next allocation on that stream without any synchronization; the device is
// only asked for memory (hipMallocAsync) on a miss.  This replaces one
// hipMallocAsync/hipFreeAsync pair per temporary buffer of every rule call.
// Each stream has its own arena and lock, so the serving workers (one stream
// each) never wait for one another here.
namespace {
struct StreamArena {
  std::mutex mu;
  std::unordered_map<size_t, std::vector<void*>> free_lists;  // size class -> blocks
  std::unordered_map<void*, size_t> live;                     // block -> size class
};
struct Arena {
  std::shared_mutex mu;  // guards the map (arenas are never removed)
  std::unordered_map<hipStream_t, std::unique_ptr<StreamArena>> streams;
};
Arena& arena() {
  static auto* a = new Arena();  // never destroyed: blocks may be released at exit
  return *a;
}
StreamArena& stream_arena(hipStream_t stream) {
  thread_local hipStream_t last_stream = nullptr;
  thread_local StreamArena* last = nullptr;
  if (last && last_stream == stream) return *last;
  Arena& a = arena();
  {
    std::shared_lock<std::shared_mutex> lock(a.mu);
    auto it = a.streams.find(stream);
    if (it != a.streams.end()) {
      last_stream = stream;
      last = it->second.get();
      return *last;
    }
  }
  std::unique_lock<std::shared_mutex> lock(a.mu);
  auto& slot = a.streams[stream];
  if (!slot) {
    slot = std::make_unique<StreamArena>();
    slot->live.reserve(1024);
  }
  last_stream = stream;
  last = slot.get();
  return *last;
}
size_t size_class(size_t bytes) {
  size_t c = 512;
  while (c < bytes) c <<= 1;
  return c;
}
// Returns a live block to the free list of the arena that allocated it.
bool recycle(StreamArena& sa, void* p) {
  std::lock_guard<std::mutex> lock(sa.mu);
  auto it = sa.live.find(p);
  if (it == sa.live.end()) return false;
  sa.free_lists[it->second].push_back(p);
  sa.live.erase(it);
  return true;
}
}  // namespace

void release_cached_blocks(hipStream_t stream) noexcept {
  StreamArena& sa = stream_arena(stream);
  std::vector<void*> blocks;
  {
    std::lock_guard<std::mutex> lock(sa.mu);
    for (auto& kv : sa.free_lists) blocks.insert(blocks.end(), kv.second.begin(), kv.second.end());
    sa.free_lists.clear();
  }
  for (void* p : blocks) (void)hipFreeAsync(p, stream);
}

void* dev_alloc(size_t bytes, hipStream_t stream, const char* context) {
  StreamArena& sa = stream_arena(stream);
  const size_t cls = size_class(bytes);
  {
    std::lock_guard<std::mutex> lock(sa.mu);
    auto f = sa.free_lists.find(cls);
    if (f != sa.free_lists.end() && !f->second.empty()) {
      void* p = f->second.back();
      f->second.pop_back();
      sa.live.emplace(p, cls);
      return p;
    }
  }
  void* p = nullptr;
  hipError_t st = hipMallocAsync(&p, cls, stream);
  if (st == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    release_cached_blocks(stream);
    st = hipMallocAsync(&p, cls, stream);
  }
  check(st, context);
  std::lock_guard<std::mutex> lock(sa.mu);
  sa.live.emplace(p, cls);
  return p;
}

void dev_free(void* p, hipStream_t stream) noexcept {
  if (!p) return;
  if (recycle(stream_arena(stream), p)) return;
  {  // allocated on another stream
    Arena& a = arena();
    StreamArena* owner = nullptr;
    {
      std::shared_lock<std::shared_mutex> lock(a.mu);
      for (auto& kv : a.streams) {
        std::lock_guard<std::mutex> l2(kv.second->mu);
        if (kv.second->live.count(p)) {
          owner = kv.second.get();
          break;
        }
      }
    }
    if (owner) {
      // The owning stream's next allocation may hand this block out with no
      // ordering against work still queued on the freeing stream: that work
      // finishes first.
      (void)hipStreamSynchronize(stream);
      if (recycle(*owner, p)) return;
    }
  }
  // Memory the caller allocated with hipMalloc (e.g. a plan's new_A).
  (void)hipStreamSynchronize(stream);
  (void)hipFree(p);
}

void row_activity(const int32_t* ptr, const int32_t* idx, const double* val, int32_t m, int32_t n, const double* l,
                  const double* u, double ztol, double* min_fin, double* max_fin, int32_t* neg_inf, int32_t* pos_inf,
                  hipStream_t stream) {
  if (m <= 0) return;
  void* lu = dev_alloc(sizeof(double) * 2 * static_cast<size_t>(n > 0 ? n : 1), stream, "packed bounds");
  hipError_t st = gpk_pack_lu(l, u, n, static_cast<double*>(lu), stream);
  if (st == hipSuccess)
    st = gpk_row_activity_lu(ptr, idx, val, m, static_cast<const double*>(lu), ztol, min_fin, max_fin, neg_inf,
                             pos_inf, stream);
  dev_free(lu, stream);  // stream-ordered reuse: later work on `stream` runs after the kernels
  check(st, "row activity");
}

void scan_mask_launch(const uint8_t* mask, int32_t n, int32_t* out) {
  Ctx& c = ctx();
  if (n <= 0) check(gpk_fill_i32(out, 0, 1, c.stream()), "scan empty");
  else check(gpk_scan_lb_mask(mask, out, n, c.lookback(gpk_lookback_tiles(n)), c.stream()), "scan mask");
}

int32_t scan_mask_count(const uint8_t* mask, int32_t n, int32_t* out) {
  scan_mask_launch(mask, n, out);
  return n <= 0 ? 0 : ctx().scalar_to_host(out + n, "scan total");
}

void scan_counts_launch(const int32_t* in, int32_t n, int32_t* out) {
  Ctx& c = ctx();
  if (n <= 0) check(gpk_fill_i32(out, 0, 1, c.stream()), "scan empty");
  else check(gpk_scan_lb_i32(in, out, n, c.lookback(gpk_lookback_tiles(n)), c.stream()), "scan counts");
}

int32_t scan_counts(const int32_t* in, int32_t n, int32_t* out) {
  scan_counts_launch(in, n, out);
  return n <= 0 ? 0 : ctx().scalar_to_host(out + n, "scan total");
}

// Terms fetched together with their count (32 KB).
constexpr int32_t kSpeculativeTerms = 4096;

double ordered_masked_sum(const uint8_t* mask, const double* vals, int32_t n) {
  if (n <= 0) return 0.0;
  Ctx& c = ctx();
  hipStream_t s = c.stream();
  DBuf<int32_t> prefix(static_cast<int64_t>(n) + 1);
  DBuf<double> packed(n);
  check(gpk_scan_lb_mask(mask, prefix, n, c.lookback(gpk_lookback_tiles(n)), s), "scan objective mask");
  check(gpk_pack_f64(mask, vals, prefix, n, packed, s), "pack objective terms");
  // The first kSpeculativeTerms packed terms travel with the count, so the
  // usual case (few terms) costs one wait.
  const int32_t head = std::min<int32_t>(n, kSpeculativeTerms);
  std::vector<double> h(static_cast<size_t>(head));
  c.download(h.data(), packed.get(), sizeof(double) * static_cast<size_t>(head), "objective terms");
  const int32_t k = c.scalar_to_host(prefix.get() + n, "objective term count");
  if (k > head) {
    h.resize(static_cast<size_t>(k));
    c.download(h.data() + head, packed.get() + head, sizeof(double) * static_cast<size_t>(k - head),
               "objective terms");
    c.sync("objective terms");
  }
  double sum = 0.0;
  for (int32_t t = 0; t < k; ++t) sum += h[static_cast<size_t>(t)];
  return sum;
}

void add_ordered_masked_sum(double* target, const uint8_t* mask, const double* vals, int32_t n) {
  Ctx& c = ctx();
  if (!c.defer_objective_sums || n <= 0) {
    *target += ordered_masked_sum(mask, vals, n);
    return;
  }
  hipStream_t s = c.stream();
  // The packed terms and the count stay allocated until the addition has run.
  struct Terms {
    DBuf<int32_t> prefix;
    DBuf<double> packed;
    std::vector<double> h;
    int32_t k = 0;
  };
  auto t = std::make_shared<Terms>();
  t->prefix.alloc(static_cast<int64_t>(n) + 1);
  t->packed.alloc(n);
  check(gpk_scan_lb_mask(mask, t->prefix, n, c.lookback(gpk_lookback_tiles(n)), s), "scan objective mask");
  check(gpk_pack_f64(mask, vals, t->prefix, n, t->packed, s), "pack objective terms");
  const int32_t head = std::min<int32_t>(n, kSpeculativeTerms);
  t->h.resize(static_cast<size_t>(head));
  c.begin_copy_group();
  c.download(t->h.data(), t->packed.get(), sizeof(double) * static_cast<size_t>(head), "objective terms");
  c.download(&t->k, t->prefix.get() + n, sizeof(int32_t), "objective term count");
  c.end_copy_group("objective terms");
  c.after_wait([t, target, head, s] {
    const int32_t k = t->k;
    if (k > head) {  // the stream is idle: the rest comes over directly
      t->h.resize(static_cast<size_t>(k));
      check(hipMemcpyAsync(t->h.data() + head, t->packed.get() + head, sizeof(double) * static_cast<size_t>(k - head),
                           hipMemcpyDeviceToHost, s),
            "objective terms");
      check(hipStreamSynchronize(s), "objective terms");
    }
    double sum = 0.0;
    for (int32_t q = 0; q < k; ++q) sum += t->h[static_cast<size_t>(q)];
    *target += sum;
  });
}

int32_t build_old_to_new(const uint8_t* mask, int32_t n, DBuf<int32_t>& map) {
  map.alloc(static_cast<int64_t>(n) + 1, "old_to_new");
  if (n == 0) return 0;
  Ctx& c = ctx();
  int32_t* count = map.get() + n;  // spare word after the map
  check(gpk_mask_to_map(mask, map, n, count, c.lookback(gpk_lookback_tiles(n)), c.stream()), "old_to_new");
  return c.scalar_to_host(count, "old_to_new count");
}

void free_csr(DeviceCsrMatrix& A) noexcept {
  hipStream_t s = ctx().stream();
  dev_free(A.rowPtr, s);
  dev_free(A.colVal, s);
  dev_free(A.nzVal, s);
  A.rowPtr = nullptr;
  A.colVal = nullptr;
  A.nzVal = nullptr;
}

// compact_csr_by_masks_gpu (gpu_presolve.cu:1881-1978): the column map in
// rank form, one mapping pass over the entries (per-entry new column and
// per-row kept counts), an exclusive scan into the row pointers, then the
// coalesced write of the kept entries.
// Up to this many 256-row tiles the write pass derives the tile offsets
// itself (gpk_compact_write_counts), one launch fewer.
constexpr int32_t kCompactFewTiles = 256;

DeviceCsrMatrix compact_csr(const DeviceCsrMatrix& A, const int32_t* row_map, const int32_t* col_map, int32_t m_new,
                            int32_t n_new, int32_t known_nnz, const gpk_rank_word* col_rank) {
  Ctx& c = ctx();
  DeviceCsrMatrix out;
  out.rows = m_new;
  out.cols = n_new;
  DBuf<int32_t> ptr(static_cast<int64_t>(m_new) + 1, "compact rowPtr");
  if (m_new == 0 || n_new == 0 || A.nnz == 0) {
    check(hipMemsetAsync(ptr.get(), 0, sizeof(int32_t) * (static_cast<size_t>(m_new) + 1), c.stream()),
          "compact empty rowPtr");
    out.rowPtr = ptr.release();
    return out;
  }
  const int32_t m_old = A.rows;
  DBuf<gpk_rank_word> rank;
  DBuf<int32_t> jn(A.nnz, "compact entry map");
  hipStream_t s = c.stream();
  if (n_new == A.cols) {
    col_rank = nullptr;  // no column dropped: every entry keeps its column index
  } else if (col_rank == nullptr) {
    rank.alloc((static_cast<int64_t>(A.cols) + 63) / 64, "compact column rank");
    check(gpk_build_rank(col_map, A.cols, rank, s), "compact column rank");
    col_rank = rank.get();
  }
  const int32_t tiles = gpk_compact_tiles(m_old);
  DBuf<int32_t> tile_off(static_cast<int64_t>(tiles) + 1, "compact tile offsets");
  check(gpk_compact_map(A.rowPtr, A.colVal, m_old, row_map, col_rank, jn, ptr, tile_off, s), "compact map");
  const bool few_tiles = tiles <= kCompactFewTiles;
  if (!few_tiles)
    check(gpk_scan_lb_i32(tile_off, tile_off, tiles, c.lookback(gpk_lookback_tiles(tiles)), s),
          "compact tile offsets");
  // Kept entries are at most A.nnz: the outputs are sized for that and the
  // kept count is read back once the entries are written.
  DBuf<int32_t> idx(A.nnz, "compact colVal");
  DBuf<double> val(A.nnz, "compact nzVal");
  if (few_tiles)
    check(gpk_compact_write_counts(A.rowPtr, A.nzVal, m_old, row_map, jn, tile_off, ptr, m_new, idx, val, s),
          "compact write");
  else
    check(gpk_compact_write(A.rowPtr, A.nzVal, m_old, row_map, jn, tile_off, ptr, m_new, idx, val, s),
          "compact write");
  out.nnz = known_nnz >= 0 ? known_nnz : c.scalar_to_host(ptr.get() + m_new, "compact kept count");
  out.rowPtr = ptr.release();
  out.colVal = idx.release();
  out.nzVal = val.release();
  return out;
}

CompactStart compact_start(const DeviceCsrMatrix& A, const int32_t* row_map, const gpk_rank_word* col_rank,
                           int32_t* kept) {
  Ctx& c = ctx();
  hipStream_t s = c.stream();
  CompactStart st;
  st.A = &A;
  st.row_map = row_map;
  st.ptr.alloc(static_cast<int64_t>(A.rows) + 1, "compact rowPtr");
  if (A.rows == 0 || A.nnz == 0) return st;  // *kept stays as the caller zeroed it
  if (col_rank == nullptr) throw std::logic_error("compact_start: the column rank words are required");
  const int32_t tiles = gpk_compact_tiles(A.rows);
  st.jn.alloc(A.nnz, "compact entry map");
  st.tile_off.alloc(static_cast<int64_t>(tiles) + 1, "compact tile offsets");
  check(gpk_compact_map_total(A.rowPtr, A.colVal, A.rows, row_map, col_rank, st.jn, st.ptr, st.tile_off, kept, s),
        "compact map");
  st.few_tiles = tiles <= kCompactFewTiles;  // the write pass derives the offsets from the raw counts
  if (!st.few_tiles)
    check(gpk_scan_lb_i32(st.tile_off, st.tile_off, tiles, c.lookback(gpk_lookback_tiles(tiles)), s),
          "compact tile offsets");
  return st;
}

DeviceCsrMatrix compact_finish(CompactStart& st, int32_t m_new, int32_t n_new, int32_t nnz_new) {
  Ctx& c = ctx();
  hipStream_t s = c.stream();
  const DeviceCsrMatrix& A = *st.A;
  DeviceCsrMatrix out;
  out.rows = m_new;
  out.cols = n_new;
  if (m_new == 0 || n_new == 0 || A.nnz == 0 || st.jn.get() == nullptr) {
    CopyBatch().zero(st.ptr.get(), static_cast<int64_t>(m_new) + 1).run("compact empty rowPtr");
    out.rowPtr = st.ptr.release();
    return out;
  }
  DBuf<int32_t> idx(A.nnz, "compact colVal");
  DBuf<double> val(A.nnz, "compact nzVal");
  if (st.few_tiles)
    check(gpk_compact_write_counts(A.rowPtr, A.nzVal, A.rows, st.row_map, st.jn, st.tile_off, st.ptr, m_new, idx, val,
                                   s),
          "compact write");
  else
    check(gpk_compact_write(A.rowPtr, A.nzVal, A.rows, st.row_map, st.jn, st.tile_off, st.ptr, m_new, idx, val, s),
          "compact write");
  out.nnz = nnz_new;
  out.rowPtr = st.ptr.release();
  out.colVal = idx.release();
  out.nzVal = val.release();
  return out;
}

namespace {
gpk_compact_side compact_side(const DeviceCsrMatrix& M, const int32_t* row_map, const gpk_rank_word* col_rank,
                              int32_t* jn, int32_t* out_ptr) {
  return gpk_compact_side{M.rowPtr, M.colVal, M.nzVal, M.rows, 0, row_map, col_rank, jn, out_ptr, nullptr, nullptr};
}
}  // namespace

CompactPairStart compact_start_pair(const DeviceCsrMatrix& A, const int32_t* row_map, const gpk_rank_word* col_rank,
                                    const DeviceCsrMatrix& AT, const int32_t* col_map, const gpk_rank_word* row_rank,
                                    int32_t* kept) {
  Ctx& c = ctx();
  hipStream_t s = c.stream();
  if (AT.rows != A.cols || AT.cols != A.rows || AT.nnz != A.nnz)
    throw std::logic_error("compact_start_pair: AT is not the transpose shape of A");
  CompactPairStart st;
  st.A = &A;
  st.AT = &AT;
  st.row_map = row_map;
  st.col_map = col_map;
  st.ptr.alloc(static_cast<int64_t>(A.rows) + 1, "compact rowPtr");
  st.t_ptr.alloc(static_cast<int64_t>(AT.rows) + 1, "compact AT rowPtr");
  if (A.rows == 0 || AT.rows == 0 || A.nnz == 0) return st;  // *kept stays as the caller zeroed it
  if (col_rank == nullptr || row_rank == nullptr)
    throw std::logic_error("compact_start_pair: both sides' rank words are required");
  const int32_t ta = gpk_compact_tiles(A.rows), tt = gpk_compact_tiles(AT.rows);
  st.jn.alloc(A.nnz, "compact entry map");
  st.t_jn.alloc(AT.nnz, "compact AT entry map");
  st.tile_off.alloc(static_cast<int64_t>(ta) + tt + 1, "compact tile offsets");
  check(gpk_compact_map_pair(compact_side(A, row_map, col_rank, st.jn, st.ptr),
                             compact_side(AT, col_map, row_rank, st.t_jn, st.t_ptr), st.tile_off, kept, s),
        "compact map pair");
  st.from_counts = ta <= kCompactFewTiles && tt <= kCompactFewTiles;
  if (!st.from_counts)
    check(gpk_scan_lb_i32(st.tile_off, st.tile_off, ta + tt, c.lookback(gpk_lookback_tiles(ta + tt)), s),
          "compact tile offsets");
  return st;
}

std::pair<DeviceCsrMatrix, DeviceCsrMatrix> compact_finish_pair(CompactPairStart& st, int32_t m_new, int32_t n_new,
                                                                int32_t nnz_new) {
  Ctx& c = ctx();
  hipStream_t s = c.stream();
  const DeviceCsrMatrix& A = *st.A;
  const DeviceCsrMatrix& AT = *st.AT;
  DeviceCsrMatrix a_out, t_out;
  a_out.rows = m_new;
  a_out.cols = n_new;
  t_out.rows = n_new;
  t_out.cols = m_new;
  if (m_new == 0 || n_new == 0 || A.nnz == 0 || st.jn.get() == nullptr) {
    CopyBatch()
        .zero(st.ptr.get(), static_cast<int64_t>(m_new) + 1)
        .zero(st.t_ptr.get(), static_cast<int64_t>(n_new) + 1)
        .run("compact empty rowPtr");
    a_out.rowPtr = st.ptr.release();
    t_out.rowPtr = st.t_ptr.release();
    return {a_out, t_out};
  }
  DBuf<int32_t> idx(A.nnz, "compact colVal"), t_idx(AT.nnz, "compact AT colVal");
  DBuf<double> val(A.nnz, "compact nzVal"), t_val(AT.nnz, "compact AT nzVal");
  gpk_compact_side a = compact_side(A, st.row_map, nullptr, st.jn, st.ptr);
  gpk_compact_side t = compact_side(AT, st.col_map, nullptr, st.t_jn, st.t_ptr);
  a.m_new = m_new;
  a.out_idx = idx;
  a.out_val = val;
  t.m_new = n_new;
  t.out_idx = t_idx;
  t.out_val = t_val;
  check(gpk_compact_write_pair(a, t, st.tile_off, st.from_counts ? 1 : 0, s), "compact write pair");
  a_out.nnz = t_out.nnz = nnz_new;
  a_out.rowPtr = st.ptr.release();
  a_out.colVal = idx.release();
  a_out.nzVal = val.release();
  t_out.rowPtr = st.t_ptr.release();
  t_out.colVal = t_idx.release();
  t_out.nzVal = t_val.release();
  return {a_out, t_out};
}

// Stable transpose: LSD radix sort of the column indices (entries already in
// row-major order, so every AT row comes out sorted by row index) carrying
// the entry positions; the row pointers come from the sorted keys (no
// per-column atomic counters), then one gather of packed (row, value) pairs
// writes rows and values.
DeviceCsrMatrix transpose_csr(const DeviceCsrMatrix& A) {
  Ctx& c = ctx();
  hipStream_t s = c.stream();
  DeviceCsrMatrix T;
  T.rows = A.cols;
  T.cols = A.rows;
  T.nnz = A.nnz;
  DBuf<int32_t> ptr(static_cast<int64_t>(A.cols) + 1, "transpose rowPtr");
  if (A.nnz == 0) {
    check(gpk_col_bounds(nullptr, 0, A.cols, ptr, s), "transpose rowPtr");
    T.rowPtr = ptr.release();
    return T;
  }
  const int32_t nnz = A.nnz;
  int32_t bits = 1;
  while (bits < 31 && (int64_t{1} << bits) < A.cols) ++bits;
  DBuf<uint32_t> ka(nnz), kb(nnz);
  DBuf<int32_t> va(nnz), vb(nnz);
  int32_t in_b = 0;
  {
    DBuf<int32_t> scratch(gpk_radix_scratch_ints(nnz));
    check(gpk_transpose_sort(A.colVal, nnz, bits, ka, va, kb, vb, scratch, &in_b, s), "transpose sort");
  }
  const uint32_t* keys = in_b ? kb.get() : ka.get();
  const int32_t* perm = in_b ? vb.get() : va.get();
  check(gpk_col_bounds(keys, nnz, A.cols, ptr, s), "transpose rowPtr");
  DBuf<gpk_row_val> rv(nnz, "transpose rows and values");
  check(gpk_expand_rows(A.rowPtr, A.nzVal, A.rows, rv, s), "expand rows");
  DBuf<int32_t> idx(nnz, "transpose colVal");
  DBuf<double> val(nnz, "transpose nzVal");
  check(gpk_transpose_gather(perm, rv, nnz, idx, val, s), "transpose gather");
  T.rowPtr = ptr.release();
  T.colVal = idx.release();
  T.nzVal = val.release();
  return T;
}

void sort_rows_if_needed(const DeviceCsrMatrix& A) {
  if (A.rows <= 0 || A.nnz <= 1) return;
  Ctx& c = ctx();
  int32_t* f = c.flags();
  check(gpk_rows_unsorted(A.rowPtr, A.colVal, A.rows, f, c.stream()), "rows unsorted check");
  if (c.read_flags(1)[0] == 0) return;
  check(gpk_sort_rows_in_place(A.rowPtr, A.colVal, A.nzVal, A.rows, c.stream()), "sort rows");
}

void check_lp_arrays(const presolve::LPInfoGpu& lp, const char* who) {
  if (lp.A.rows <= 0 && lp.AT.rows <= 0) return;
  Ctx& c = ctx();
  int32_t* f = c.flags();
  check(gpk_csr_check(lp.A.rowPtr, lp.A.colVal, lp.A.rows, lp.A.cols, lp.A.nnz, f, c.stream()), "check A");
  check(gpk_csr_check(lp.AT.rowPtr, lp.AT.colVal, lp.AT.rows, lp.AT.cols, lp.AT.nnz, f + 4, c.stream()), "check AT");
  int32_t h[8];
  std::memcpy(h, c.read_flags(8), sizeof(h));
  const bool a_again = h[0] == 0 && h[2] == 0 && h[3] != 0, at_again = h[4] == 0 && h[6] == 0 && h[7] != 0;
  if (a_again || at_again) {  // unsorted rows: look for columns repeated at a distance
    f = c.flags();
    if (a_again) check(gpk_csr_check_rows(lp.A.rowPtr, lp.A.colVal, lp.A.rows, lp.A.cols, lp.A.nnz, f, c.stream()),
                       "check A rows");
    if (at_again)
      check(gpk_csr_check_rows(lp.AT.rowPtr, lp.AT.colVal, lp.AT.rows, lp.AT.cols, lp.AT.nnz, f + 4, c.stream()),
            "check AT rows");
    const int32_t* g = c.read_flags(8);
    for (int k = 0; k < 8; ++k) h[k] |= g[k];
  }
  static const char* const what[3] = {"row pointers are not monotone from 0 to nnz",
                                      "a column index is outside [0, cols)", "a row repeats a column index"};
  for (int m = 0; m < 2; ++m)
    for (int k = 0; k < 3; ++k)
      if (h[4 * m + k])
        throw InputError(std::string(who) + ": " + (m == 0 ? "A" : "AT") + ": " + what[k]);
}

void append_record(PostsolveTape& tape, PostsolveReductionType type, const int32_t* idx, size_t ni,
                   const double* vals, size_t nv, PostsolveDualMode mode) {
  tape.types.push_back(static_cast<int32_t>(type));
  tape.indices.insert(tape.indices.end(), idx, idx + ni);
  tape.vals.insert(tape.vals.end(), vals, vals + nv);
  tape.index_starts.push_back(static_cast<int32_t>(tape.indices.size()));
  tape.value_starts.push_back(static_cast<int32_t>(tape.vals.size()));
  tape.dual_modes.push_back(static_cast<uint8_t>(mode));
}

double fixed_col_tape(PostsolveTape* tape, const uint8_t* fixed_mask, const uint8_t* excluded,
                      const double* fixed_val, const uint8_t* keep_row, const double* cvec, const DeviceCsrMatrix& AT,
                      int32_t n) {
  if (n <= 0) return 0.0;
  Ctx& c = ctx();
  hipStream_t s = c.stream();
  const gpk_csr at = csr_view(AT);
  DBuf<uint8_t> sel(n);
  DBuf<int32_t> sel_prefix(static_cast<int64_t>(n) + 1), nnz_prefix(static_cast<int64_t>(n) + 1);
  check(gpk_fixed_col_counts(at, fixed_mask, excluded, keep_row, sel, nnz_prefix, s), "fixed col counts");
  scan_mask_launch(sel, n, sel_prefix);
  scan_counts_launch(nnz_prefix, n, nnz_prefix);
  int32_t records = 0, active = 0;  // both totals in one wait
  c.begin_copy_group();
  c.download(&records, sel_prefix.get() + n, sizeof(int32_t), "fixed col records");
  c.download(&active, nnz_prefix.get() + n, sizeof(int32_t), "fixed col entries");
  c.sync("fixed col counts");
  if (records == 0) return 0.0;
  DBuf<int32_t> cols(records), rows(active), rstart(records);
  DBuf<double> v(records), cc(records), coef(active);
  check(gpk_fixed_col_pack(at, sel, sel_prefix, nnz_prefix, fixed_val, cvec, keep_row, cols, v, cc, rstart, rows, coef,
                           s),
        "fixed col pack");
  // Everything the objective and the tape need, in one wait.
  std::vector<double> hv, hc, hcoef;
  std::vector<int32_t> hcol, hrow, starts;
  c.begin_copy_group();
  c.to_host_async(hv, v.get(), records, "fixed col values");
  c.to_host_async(hc, cc.get(), records, "fixed col costs");
  if (tape) {
    c.to_host_async(hcol, cols.get(), records, "fixed col ids");
    c.to_host_async(hrow, rows.get(), active, "fixed col rows");
    c.to_host_async(hcoef, coef.get(), active, "fixed col coefs");
    c.to_host_async(starts, rstart.get(), records, "fixed col starts");
  }
  c.sync("fixed col tape");
  long double obj = 0.0L;
  for (int32_t r = 0; r < records; ++r)
    obj += static_cast<long double>(hc[r]) * static_cast<long double>(hv[r]);
  if (tape) {
    starts.push_back(active);
    std::vector<int32_t> idx;
    std::vector<double> vals;
    for (int32_t r = 0; r < records; ++r) {
      idx.assign(1, hcol[r]);
      vals.assign({hv[r], hc[r]});
      for (int32_t q = starts[r]; q < starts[r + 1]; ++q) {
        idx.push_back(hrow[q]);
        vals.push_back(hcoef[q]);
      }
      append_record(*tape, PostsolveReductionType::FixedCol, idx.data(), idx.size(), vals.data(), vals.size(),
                    PostsolveDualMode::Minimal);
    }
  }
  return static_cast<double>(obj);
}

}  // namespace gpu_presolver::amd
