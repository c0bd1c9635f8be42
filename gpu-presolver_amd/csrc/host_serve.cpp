// Serving runtime (no counterpart in the reference, which presolves one LP
// per call on the legacy stream): a FIFO job queue, admission by a device
// memory budget, and a pool of worker threads.  Each worker owns an
// execution context (its own non-blocking stream, flag ring and pool free
// lists, host_common.cpp), so the jobs in flight overlap on the device while
// each one runs exactly the kernels it would run alone: every result is
// bit-identical to the same job processed by itself.
//
// A job: presolve of a caller-owned device LP; then, when postsolve is on, a
// seeded reduced-space point in place of an LP solver and postsolve against
// the original LP; the postsolved x, y, z are reduced to an order-free
// checksum of their bit patterns and returned to the pool.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "gp_host.hpp"
#include "gp_presolve_capi.h"
#include "gpu_presolver/presolve/gpu_postsolve.hpp"
#include "gpu_presolver/presolve/gpu_presolve.hpp"

namespace gpu_presolver::amd {

using presolve::GpuPostsolveResult;
using presolve::GpuPresolveSummary;
using presolve::LPInfoGpu;
using presolve::PresolveParams;

namespace {

// Device bytes of an LP (A, AT, five vectors).
double lp_bytes(const LPInfoGpu& lp) {
  const double m = lp.A.rows, n = lp.A.cols, nnz = lp.A.nnz;
  return 4.0 * (m + 1) + 4.0 * (n + 1) + 24.0 * nnz + 8.0 * (3.0 * n + 2.0 * m);
}

// Device bytes of the LP of a spec (expected entries per general row).
double spec_bytes(const gpk_lpgen_spec& sp) {
  const double len = sp.mean_len > 0.0 ? sp.mean_len : 0.5 * (sp.len_lo + sp.len_hi);
  const double m = sp.m, n = sp.n, nnz = m * (0.9 * len + 0.15) + 0.05 * n;
  return 4.0 * (m + 1) + 4.0 * (n + 1) + 24.0 * nnz + 8.0 * (3.0 * n + 2.0 * m);
}

}  // namespace

void check_spec(const gpk_lpgen_spec& sp) {
  if (sp.m < 1 || sp.n < 1) throw InputError("gp_lp_spec: m and n must be >= 1");
  if (sp.mean_len > 0.0) {
    if (!(sp.mean_len <= 64.0) || sp.min_len < 0 || sp.min_len > 64)
      throw InputError("gp_lp_spec: mean_len must be in (0, 64] and min_len in [0, 64]");
  } else if (sp.len_lo < 0 || sp.len_hi < sp.len_lo || sp.len_hi > 256) {
    throw InputError("gp_lp_spec: need 0 <= len_lo <= len_hi <= 256");
  }
}

// Generated LP in blocks of the calling thread's pool (k_lpgen.hip); two
// waits (the raw and the final entry counts).
LPInfoGpu generate_lp(const gpk_lpgen_spec& sp) {
  check_spec(sp);
  Ctx& c = ctx();
  hipStream_t s = c.stream();
  const int32_t m = sp.m, n = sp.n;
  DBuf<uint8_t> ccls(n), rcls(m), kind(m);
  DBuf<double> x0(n), cv(n), lv(n), uv(n), AL(m), AU(m);
  check(gpk_lpgen_columns(sp, ccls, cv, lv, uv, x0, s), "generate columns");
  DBuf<int32_t> raw_count(static_cast<int64_t>(m) + 1), scol_row(n), extra(m), own(m), cursor(m);
  CopyBatch().zero(extra.get(), m).zero(cursor.get(), m).run("generator counters");
  check(gpk_lpgen_row_counts(sp, ccls, rcls, kind, raw_count, scol_row, extra, s), "generate rows");
  DBuf<int32_t> raw_ptr(static_cast<int64_t>(m) + 1);
  const int32_t raw_nnz = scan_counts(raw_count, m, raw_ptr);
  DBuf<int32_t> raw_col(raw_nnz), unique(m), src(m), len(static_cast<int64_t>(m) + 1);
  DBuf<double> raw_val(raw_nnz);
  check(gpk_lpgen_fill(sp, ccls, rcls, raw_ptr, scol_row, own, cursor, raw_col, raw_val, unique, src, len, s),
        "generate entries");
  DBuf<int32_t> ptr(static_cast<int64_t>(m) + 1);
  const int32_t nnz = scan_counts(len, m, ptr);
  DBuf<int32_t> idx(nnz);
  DBuf<double> val(nnz);
  check(gpk_lpgen_write(sp, rcls, kind, raw_ptr, raw_col, raw_val, src, ptr, idx, val, x0, AL, AU, s),
        "generate matrix");
  LPInfoGpu lp;
  lp.A = presolve::DeviceCsrMatrix{m, n, nnz, ptr.get(), idx.get(), val.get()};
  lp.AT = transpose_csr(lp.A);
  (void)ptr.release();
  (void)idx.release();
  (void)val.release();
  lp.c = cv.release();
  lp.l = lv.release();
  lp.u = uv.release();
  lp.AL = AL.release();
  lp.AU = AU.release();
  lp.obj_constant = 0.0;
  return lp;
}

void free_generated_lp(LPInfoGpu& lp) noexcept {
  hipStream_t s = ctx().stream();
  free_csr(lp.A);
  free_csr(lp.AT);
  for (double** v : {&lp.c, &lp.l, &lp.u, &lp.AL, &lp.AU}) {
    dev_free(*v, s);
    *v = nullptr;
  }
}

// Salts of the three vectors in gp_checksum.
constexpr uint64_t kSaltX = 0, kSaltY = 1ull << 40, kSaltZ = 2ull << 40;

uint64_t checksum_xyz(const double* x, int32_t n, const double* y, int32_t m, const double* z) {
  Ctx& c = ctx();
  auto* acc = reinterpret_cast<unsigned long long*>(c.flags());  // two zeroed words, 64-byte aligned
  check(gpk_checksum_f64(x, n, kSaltX, acc, c.stream()), "checksum x");
  check(gpk_checksum_f64(y, m, kSaltY, acc, c.stream()), "checksum y");
  check(gpk_checksum_f64(z, n, kSaltZ, acc, c.stream()), "checksum z");
  const int32_t* w = c.read_flags(2);
  return static_cast<uint64_t>(static_cast<uint32_t>(w[0])) | (static_cast<uint64_t>(static_cast<uint32_t>(w[1])) << 32);
}

// One job on the calling thread's context: presolve through the public
// entry point; with postsolve, the seeded reduced point, the postsolve
// against the original LP and the checksum of x, y, z.  The serving workers
// run it in job mode (own stream, pooled outputs); gp_plain_job runs it in
// the default mode of the public API (device synchronization on entry,
// hipMalloc'ed outputs), i.e. exactly one-at-a-time plain processing.
gp_job_result run_job(const LPInfoGpu& lp, const PresolveParams& prm, bool postsolve, uint64_t seed, uint64_t tag) {
  gp_job_result r{};
  r.tag = tag;
  Ctx& c = ctx();
  try {
    AbandonOnThrow abandon_guard;
    GpuPresolveSummary sum = presolve::run_gpu_presolve_with_reduced_lp(lp, prm);
    r.reduced_rows = sum.reduced_rows;
    r.reduced_cols = sum.reduced_cols;
    r.reduced_nnz = sum.reduced_nnz;
    r.iterations = sum.iterations;
    r.has_infeasible = sum.has_infeasible;
    r.has_unbounded = sum.has_unbounded;
    try {
      if (postsolve && !sum.has_infeasible && !sum.has_unbounded) {
        const LPInfoGpu& red = sum.reduced_lp;
        const int32_t n1 = red.A.cols, m1 = red.A.rows;
        DBuf<double> x(n1), y(m1), z(n1);
        check(gpk_synthetic_point(seed ^ tag, n1, m1, red.l, red.u, x, y, z, c.stream()), "synthetic reduced point");
        GpuPostsolveResult ps = presolve::postsolve_gpu(x, y, z, sum.record, &lp, nullptr);
        try {
          r.checksum = checksum_xyz(ps.x_org, ps.n0, ps.y_org, ps.m0, ps.z_org);
        } catch (...) {
          dev_free(ps.x_org, c.stream());
          dev_free(ps.y_org, c.stream());
          dev_free(ps.z_org, c.stream());
          throw;
        }
        dev_free(ps.x_org, c.stream());
        dev_free(ps.y_org, c.stream());
        dev_free(ps.z_org, c.stream());
      }
    } catch (...) {
      presolve::free_gpu_presolve_reduced_lp(sum);
      throw;
    }
    presolve::free_gpu_presolve_reduced_lp(sum);
    c.sync("job");
    r.status = GP_OK;
  } catch (const HipError&) {
    r.status = GP_ERR_DEVICE;
  } catch (const std::bad_alloc&) {
    r.status = GP_ERR_INTERNAL;
  } catch (const std::exception&) {
    r.status = GP_ERR_INTERNAL;
  }
  return r;
}

// A job on a generated LP: generation on the calling thread's stream, then
// run_job; `ready` is called once the LP is resident.
template <class F>
gp_job_result run_generated_job(const gpk_lpgen_spec& sp, const PresolveParams& prm, bool postsolve, uint64_t seed,
                                uint64_t tag, F&& ready) {
  LPInfoGpu lp;
  try {
    AbandonOnThrow abandon_guard;
    lp = generate_lp(sp);
    ctx().sync("generated LP");
  } catch (const HipError&) {
    free_generated_lp(lp);
    gp_job_result r{};
    r.tag = tag;
    r.status = GP_ERR_DEVICE;
    return r;
  } catch (const std::exception&) {
    free_generated_lp(lp);
    gp_job_result r{};
    r.tag = tag;
    r.status = GP_ERR_INTERNAL;
    return r;
  }
  ready();
  gp_job_result r = run_job(lp, prm, postsolve, seed, tag);
  free_generated_lp(lp);
  return r;
}

gp_job_result run_generated_job(const gpk_lpgen_spec& sp, const PresolveParams& prm, bool postsolve, uint64_t seed,
                                uint64_t tag, double* t_ready) {
  return run_generated_job(sp, prm, postsolve, seed, tag, [&] {
    *t_ready = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  });
}

class Server {
 public:
  struct Job {
    int64_t id;
    LPInfoGpu lp;
    uint64_t tag;
    double t_arrival;
    double bytes;
    bool generate;  // the worker generates the LP from spec first
    gpk_lpgen_spec spec;
  };

  Server(const PresolveParams& prm, const gp_serve_config& cfg) : prm_(prm), cfg_(cfg) {
    if (cfg_.workers < 1) throw std::invalid_argument("gp_serve_config.workers must be >= 1");
    check(hipGetDevice(&device_), "hipGetDevice");
    t0_ = std::chrono::steady_clock::now();
    for (int w = 0; w < cfg_.workers; ++w) ctxs_.push_back(std::make_unique<Ctx>());
    for (int w = 0; w < cfg_.workers; ++w) workers_.emplace_back([this, w] { worker_main(w); });
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> lock(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    for (auto& t : workers_) t.join();
    ctxs_.clear();
    if (std::getenv("GP_SERVE_PROFILE") != nullptr && jobs_run_ > 0)
      std::fprintf(stderr, "serving profile: %lld jobs, per job %.3f ms on a worker, of which %.3f ms waiting for "
                   "the GPU\n", static_cast<long long>(jobs_run_), 1e3 * job_s_ / jobs_run_,
                   1e3 * wait_s_ / jobs_run_);
  }

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
  }

  int64_t submit(const LPInfoGpu& lp, uint64_t tag, double t_arrival) {
    Job j{0, lp, tag, t_arrival < 0.0 ? now() : t_arrival, 4.0 * lp_bytes(lp), false, {}};
    return enqueue(j);
  }

  int64_t submit_spec(const gpk_lpgen_spec& sp, uint64_t tag, double t_arrival) {
    check_spec(sp);
    // the generated LP and its generation scratch count with the job
    Job j{0, LPInfoGpu{}, tag, t_arrival < 0.0 ? now() : t_arrival, 6.0 * spec_bytes(sp), true, sp};
    return enqueue(j);
  }

  int64_t enqueue(Job j) {
    if (j.bytes <= 0.0) j.bytes = 1.0;
    {
      std::lock_guard<std::mutex> lock(mu_);
      j.id = next_id_++;
      queue_.push_back(j);
    }
    cv_work_.notify_one();
    return j.id;
  }

  int wait(gp_job_result* out, int max, double timeout_s) {
    std::unique_lock<std::mutex> lock(mu_);
    const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(std::max(0.0, timeout_s));
    cv_done_.wait_until(lock, until, [&] { return !done_.empty(); });
    int k = 0;
    while (k < max && !done_.empty()) {
      out[k++] = done_.front();
      done_.pop_front();
    }
    return k;
  }

 private:

  int64_t jobs_run_ = 0;  // worker-side totals for GP_SERVE_PROFILE
  double job_s_ = 0.0, wait_s_ = 0.0;

  // FIFO admission: the head job starts when a worker is free and it fits
  // the budget next to the jobs in flight (or nothing else is in flight).
  bool admissible() const {
    if (queue_.empty()) return false;
    if (inflight_jobs_ == 0 || cfg_.hbm_budget_bytes <= 0.0) return true;
    return inflight_bytes_ + queue_.front().bytes <= cfg_.hbm_budget_bytes;
  }

  void worker_main(int w) {
    (void)hipSetDevice(device_);
    ScopedCtx scope(*ctxs_[w]);
    job_mode() = JobMode{false, true};
    while (true) {
      Job j;
      {
        std::unique_lock<std::mutex> lock(mu_);
        cv_work_.wait(lock, [&] { return stop_ && queue_.empty() ? true : admissible(); });
        if (queue_.empty()) return;  // stopping and drained
        j = queue_.front();
        queue_.pop_front();
        inflight_bytes_ += j.bytes;
        ++inflight_jobs_;
      }
      const double w0 = ctx().wait_seconds();
      const auto t0 = std::chrono::steady_clock::now();
      gp_job_result r = run(j, w);
      const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      const double dw = ctx().wait_seconds() - w0;
      {
        std::lock_guard<std::mutex> lock(mu_);
        ++jobs_run_;
        job_s_ += dt;
        wait_s_ += dw;
        inflight_bytes_ -= j.bytes;
        --inflight_jobs_;
        done_.push_back(r);
      }
      cv_done_.notify_all();
      cv_work_.notify_all();
    }
  }

  gp_job_result run(const Job& j, int w) {
    const double t_start = now();
    gp_job_result r{};
    double t_ready = t_start;
    if (!j.generate) {
      r = run_job(j.lp, prm_, cfg_.postsolve != 0, cfg_.seed, j.tag);
    } else {
      r = run_generated_job(j.spec, prm_, cfg_.postsolve != 0, cfg_.seed, j.tag,
                            [&] { t_ready = now(); });
    }
    r.job = j.id;
    r.tag = j.tag;
    r.worker = w;
    r.t_arrival = j.t_arrival;
    r.t_start = t_start;
    r.t_ready = t_ready;
    r.t_end = now();
    return r;
  }

  PresolveParams prm_;
  gp_serve_config cfg_;
  int device_ = 0;
  std::chrono::steady_clock::time_point t0_;
  std::vector<std::unique_ptr<Ctx>> ctxs_;
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Job> queue_;
  std::deque<gp_job_result> done_;
  double inflight_bytes_ = 0.0;
  int inflight_jobs_ = 0;
  int64_t next_id_ = 0;
  bool stop_ = false;
};

Server* server_create(const PresolveParams& prm, const gp_serve_config& cfg) { return new Server(prm, cfg); }
void server_destroy(Server* s) { delete s; }
int64_t server_submit(Server* s, const LPInfoGpu& lp, uint64_t tag, double t_arrival) {
  return s->submit(lp, tag, t_arrival);
}
int64_t server_submit_spec(Server* s, const gpk_lpgen_spec& sp, uint64_t tag, double t_arrival) {
  return s->submit_spec(sp, tag, t_arrival);
}
int server_wait(Server* s, gp_job_result* out, int max, double timeout_s) { return s->wait(out, max, timeout_s); }
double server_now(const Server* s) { return s->now(); }

}  // namespace gpu_presolver::amd
