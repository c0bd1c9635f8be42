// Restarted primal-dual hybrid gradient (PDHG) for the reduced LP
//   min c^T x + k  s.t.  AL <= A x <= AU,  l <= x <= u,
// run on the LP that presolve leaves in device memory, so the reduced LP
// never leaves the GPU between presolve, solve and postsolve.  Method 0 is
// the published restarted-average PDHG of Applegate et al. ("Practical
// large-scale linear programming using primal-dual hybrid gradient",
// NeurIPS 2021): Ruiz equilibration, step sizes from a power-iteration
// estimate of ||A||_2, restarts to the better of the current and the
// average iterate by the KKT error (sufficient 0.2, necessary 0.8,
// artificial 0.36 of the iterations so far), and the primal weight
// rebalanced at every restart from the distance travelled.  Method 1
// replaces the averaging by the reflected Halpern iteration of Lu and Yang
// ("Restarted Halpern PDHG for linear programming", 2024), the
// Halpern-Peaceman-Rachford form: z <- (k+1)/(k+2) (2 T(z) - z) +
// 1/(k+2) z_restart with T the PDHG step; the KKT error is taken at T(z),
// which is also the restart point and the returned point.
//
// Each check period of `check_every` iterations is one graph launch (or the
// same kernels launched directly); tau and sigma live in device memory so
// the captured iterations read the current step sizes.  The KKT sums of the
// current and the average iterate are reduced on the device and read back
// with one wait per period.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>

#include "gp_host.hpp"
#include "gp_kernels.h"

namespace gpu_presolver::amd {

using presolve::LPInfoGpu;

namespace {

gpk_csr view(const presolve::DeviceCsrMatrix& M, const double* val) {
  return gpk_csr{M.rows, M.cols, M.nnz, M.rowPtr, M.colVal, val};
}

struct Kkt {
  double pres = 0.0, dres = 0.0, pobj = 0.0, dobj = 0.0, gap = 0.0;  // absolute
  double rel_p = 0.0, rel_d = 0.0, rel_gap = 0.0;
  bool optimal(double tol) const { return rel_p <= tol && rel_d <= tol && rel_gap <= tol; }
  double weighted(double w) const { return std::sqrt(w * w * pres * pres + dres * dres / (w * w) + gap * gap); }
};

// r: row sums [0..3] then column sums [4..7] (gpk_pdhg_row_terms / col_terms).
Kkt kkt_from(const double* r, double k) {
  Kkt e;
  e.pres = std::sqrt(std::max(r[0], 0.0));
  e.dres = std::sqrt(std::max(r[4] + r[3], 0.0));
  e.pobj = r[7] + k;
  e.dobj = r[1] + r[5] + k;
  e.gap = std::fabs(e.pobj - e.dobj);
  e.rel_p = e.pres / (1.0 + std::sqrt(std::max(r[2], 0.0)));
  e.rel_d = e.dres / (1.0 + std::sqrt(std::max(r[6], 0.0)));
  e.rel_gap = e.gap / (1.0 + std::fabs(e.pobj) + std::fabs(e.dobj));
  return e;
}

struct GraphExec {
  hipGraphExec_t exec = nullptr;
  ~GraphExec() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

}  // namespace

PdhgReport pdhg_solve(const LPInfoGpu& lp, const PdhgOptions& opt, double* x_out, double* y_out, double* z_out) {
  const auto t0 = std::chrono::steady_clock::now();
  const auto seconds = [&] {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  const int32_t m = lp.A.rows, n = lp.A.cols, nnz = lp.A.nnz;
  if (opt.check_every < 1 || opt.max_iters < 0 || opt.ruiz_iters < 0 || opt.power_iters < 1 || !(opt.tol > 0.0) ||
      (opt.method != 0 && opt.method != 1))
    throw InputError(
        "pdhg: check_every, power_iters >= 1, max_iters, ruiz_iters >= 0, tol > 0 and method 0 or 1 are required");
  if (n > 0 && x_out == nullptr) throw InputError("pdhg: null x");
  if (m > 0 && y_out == nullptr) throw InputError("pdhg: null y");
  if (n > 0 && z_out == nullptr) throw InputError("pdhg: null z");
  Ctx& cx = ctx();
  const hipStream_t s = cx.stream();
  const char* where = "pdhg";
  check_lp_arrays(lp, "pdhg input");

  // ---- row blocks of A and AT (from the host copies of the row pointers) ---
  const auto row_blocks = [&](const presolve::DeviceCsrMatrix& M, DBuf<int32_t>& blk) {
    const std::vector<int32_t> ptr = cx.to_host(M.rowPtr, static_cast<int64_t>(M.rows) + 1, where);
    std::vector<int32_t> b(static_cast<size_t>(M.rows) + 1);
    const int32_t nb = gpk_pdhg_row_blocks(ptr.data(), M.rows, b.data());
    blk.alloc(static_cast<int64_t>(nb) + 1, where);
    cx.to_device(blk.get(), b.data(), static_cast<int64_t>(nb) + 1, where);
    return nb;
  };
  DBuf<int32_t> blk_a, blk_at;
  const int32_t nb_a = row_blocks(lp.A, blk_a), nb_at = row_blocks(lp.AT, blk_at);

  // ---- Ruiz equilibration: As = Dr A Dc, ATs = Dc AT Dr ------------------
  DBuf<double> As(nnz, where), ATs(nnz, where), dr(m, where), dc(n, where);
  check(gpk_fill_f64(dr, 1.0, m, s), where);
  check(gpk_fill_f64(dc, 1.0, n, s), where);
  CopyBatch().copy(As.get(), lp.A.nzVal, nnz).copy(ATs.get(), lp.AT.nzVal, nnz).run(where);
  const gpk_rows A0{view(lp.A, lp.A.nzVal), blk_a, nb_a}, AT0{view(lp.AT, lp.AT.nzVal), blk_at, nb_at};
  const gpk_rows A{view(lp.A, As), blk_a, nb_a}, AT{view(lp.AT, ATs), blk_at, nb_at};
  for (int32_t k = 0; k < opt.ruiz_iters && nnz > 0; ++k) {
    check(gpk_pdhg_ruiz(A, dr, s), where);
    check(gpk_pdhg_ruiz(AT, dc, s), where);
    check(gpk_pdhg_scale_csr(A0, dr, dc, As, s), where);
    check(gpk_pdhg_scale_csr(AT0, dc, dr, ATs, s), where);
  }
  DBuf<double> cs(n, where), ls(n, where), us(n, where), ALs(m, where), AUs(m, where);
  check(gpk_pdhg_scale_vec(lp.c, dc, cs, n, 1, s), where);
  check(gpk_pdhg_scale_vec(lp.l, dc, ls, n, 0, s), where);
  check(gpk_pdhg_scale_vec(lp.u, dc, us, n, 0, s), where);
  check(gpk_pdhg_scale_vec(lp.AL, dr, ALs, m, 1, s), where);
  check(gpk_pdhg_scale_vec(lp.AU, dr, AUs, m, 1, s), where);

  // ---- work vectors --------------------------------------------------------
  DBuf<double> x(n, where), xbar(n, where), xsum(n, where), xr(n, where), xa(n, where), xo(n, where);
  DBuf<double> y(m, where), ysum(m, where), yr(m, where), ya(m, where), yo(m, where);
  DBuf<double> partial(std::max({gpk_pdhg_partials(nb_a), gpk_pdhg_partials(nb_at),
                                 gpk_pdhg_partials(static_cast<int32_t>((std::max(m, n) + 255) / 256))}),
                       where);
  DBuf<double> res(24, where), step(3, where);
  check(gpk_pdhg_project_zero(ls, us, x, n, s), where);
  CopyBatch().zero(y.get(), m).zero(xsum.get(), n).zero(ysum.get(), m).copy(xr.get(), x.get(), n).zero(yr.get(), m)
      .zero(xbar.get(), n).run(where);

  // ---- ||As||_2 by power iteration on As^T As (normalised on the device;
  // one wait at the end) ------------------------------------------------------
  double norm = 0.0;
  if (nnz > 0) {
    DBuf<double> v(n, where), w(m, where);
    check(gpk_pdhg_seed(v, n, s), where);
    for (int32_t k = 0; k < opt.power_iters; ++k) {
      check(gpk_pdhg_spmv(A, v, w, s), where);
      check(gpk_pdhg_spmv(AT, w, v, s), where);
      check(gpk_pdhg_sq_terms(v, nullptr, n, partial, res, s), where);
      check(gpk_pdhg_normalize(v, res, n, s), where);
    }
    // ||As^T As v|| with ||v|| = 1 tends to ||As||^2
    norm = std::sqrt(std::sqrt(cx.to_host(res.get(), 1, where)[0]));
  }
  if (!(norm > 0.0)) norm = 1.0;
  const double eta = 0.998 / norm;

  // Scaled norms of c and of the row bounds set the first primal weight.
  check(gpk_pdhg_row_terms(A, ALs, AUs, x, y, partial, res.get(), s), where);
  check(gpk_pdhg_col_terms(AT, cs, ls, us, x, y, nullptr, partial, res.get() + 4, s), where);
  std::vector<double> h = cx.to_host(res.get(), 8, where);
  const double nc = std::sqrt(h[6]), nb = std::sqrt(h[2]);
  double omega = nc > 1e-10 && nb > 1e-10 ? nc / nb : 1.0;
  // step = {tau, sigma, iterations since the restart (the Halpern weights)}
  int32_t since = 0;
  const auto set_steps = [&] {
    const double st[3] = {eta / omega, eta * omega, static_cast<double>(since)};
    cx.upload(step.get(), st, sizeof st, where);
  };
  set_steps();
  const bool halpern = opt.method == 1;

  // ---- KKT of an iterate (scaled xv, yv) into res[off .. off + 8) ---------
  const auto enqueue_kkt = [&](const double* xv, const double* yv, double* z, int off) {
    check(gpk_pdhg_unscale_x(xv, dc, lp.l, lp.u, xo, n, s), where);
    check(gpk_pdhg_mul(yv, dr, yo, m, s), where);
    check(gpk_pdhg_row_terms(A0, lp.AL, lp.AU, xo, yo, partial, res.get() + off, s), where);
    check(gpk_pdhg_col_terms(AT0, lp.c, lp.l, lp.u, xo, yo, z, partial, res.get() + off + 4, s), where);
  };

  // ---- a check period: captured once, replayed -----------------------------
  const int32_t K = opt.check_every;
  const auto enqueue_period = [&] {
    return gpk_pdhg_iterations(A, AT, cs, ls, us, ALs, AUs, x, xbar, xsum, y, ysum, xr, yr, step, halpern ? 1 : 0, K,
                               s);
  };
  GraphExec graph;
  if (opt.use_graph) {
    cx.sync("pdhg graph capture");
    check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "pdhg graph capture");
    const hipError_t launched = enqueue_period();
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(s, &g);
    if ((launched != hipSuccess || end != hipSuccess) && g != nullptr) (void)hipGraphDestroy(g);
    check(launched, "pdhg graph capture");
    check(end, "pdhg graph capture");
    const hipError_t inst = hipGraphInstantiate(&graph.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    check(inst, "pdhg graph instantiate");
  }

  PdhgReport rep;
  rep.norm_estimate = norm;
  enqueue_kkt(x, y, nullptr, 0);
  h = cx.to_host(res.get(), 8, where);
  Kkt last_restart = kkt_from(h.data(), lp.obj_constant);
  double prev_candidate = std::numeric_limits<double>::infinity();
  // Which buffers hold the candidate point: the iterate (x, y), the
  // average (xa, ya) or, for Halpern, the PDHG point T(z) in (xsum, ysum).
  enum class Point { Iterate, Average, Pdhg } point = Point::Iterate;
  rep.status = last_restart.optimal(opt.tol) ? 0 : 1;

  while (rep.status != 0 && rep.iterations < opt.max_iters) {
    if (seconds() > opt.max_time) {
      rep.status = 2;
      break;
    }
    if (halpern && since > 0) set_steps();  // the period's first Halpern count
    if (graph.exec) {
      check(hipGraphLaunch(graph.exec, s), "pdhg graph launch");
      ++rep.graph_launches;
    } else {
      check(enqueue_period(), where);
    }
    rep.iterations += K;
    since += K;
    ++rep.checks;
    Kkt cand;
    if (halpern) {
      enqueue_kkt(xsum, ysum, nullptr, 0);
      h = cx.to_host(res.get(), 8, where);
      cand = kkt_from(h.data(), lp.obj_constant);
      point = Point::Pdhg;
    } else {
      check(gpk_pdhg_axpby(1.0 / since, xsum, 0.0, nullptr, xa, n, s), where);
      check(gpk_pdhg_axpby(1.0 / since, ysum, 0.0, nullptr, ya, m, s), where);
      enqueue_kkt(x, y, nullptr, 0);
      enqueue_kkt(xa, ya, nullptr, 8);
      h = cx.to_host(res.get(), 16, where);
      const Kkt cur = kkt_from(h.data(), lp.obj_constant), avg = kkt_from(h.data() + 8, lp.obj_constant);
      const bool opt_cur = cur.optimal(opt.tol), opt_avg = avg.optimal(opt.tol);
      const bool take_avg = opt_cur || opt_avg ? !opt_cur : avg.weighted(omega) < cur.weighted(omega);
      point = take_avg ? Point::Average : Point::Iterate;
      cand = take_avg ? avg : cur;
    }
    if (opt.verbose)
      std::fprintf(stderr, "pdhg it %d %s pres %.3e dres %.3e gap %.3e obj %.10g omega %.3e\n", rep.iterations,
                   point == Point::Average ? "avg" : point == Point::Pdhg ? "T(z)" : "cur", cand.rel_p, cand.rel_d,
                   cand.rel_gap, cand.pobj, omega);
    if (cand.optimal(opt.tol)) {
      rep.status = 0;
      break;
    }
    const double e = cand.weighted(omega), e0 = last_restart.weighted(omega);
    const bool restart = e <= 0.2 * e0 || (e <= 0.8 * e0 && e > prev_candidate) || since >= 0.36 * rep.iterations;
    if (!restart) {
      prev_candidate = e;
      continue;
    }
    if (point == Point::Average) CopyBatch().copy(x.get(), xa.get(), n).copy(y.get(), ya.get(), m).run(where);
    if (point == Point::Pdhg) CopyBatch().copy(x.get(), xsum.get(), n).copy(y.get(), ysum.get(), m).run(where);
    check(gpk_pdhg_sq_terms(x, xr, n, partial, res.get() + 16, s), where);
    check(gpk_pdhg_sq_terms(y, yr, m, partial, res.get() + 20, s), where);
    h = cx.to_host(res.get() + 16, 8, where);
    const double dx = std::sqrt(h[0]), dy = std::sqrt(h[4]);
    if (dx > 1e-10 && dy > 1e-10) omega = std::exp(0.5 * std::log(dy / dx) + 0.5 * std::log(omega));
    since = 0;
    set_steps();
    // (xr, yr): the restart point, also the Halpern anchor of the next run
    CopyBatch().copy(xr.get(), x.get(), n).copy(yr.get(), y.get(), m).zero(xsum.get(), n).zero(ysum.get(), m)
        .run(where);
    last_restart = cand;
    prev_candidate = std::numeric_limits<double>::infinity();
    point = Point::Iterate;  // x, y now hold the restart point
    ++rep.restarts;
  }

  // ---- the chosen iterate, unscaled, with its projected reduced costs -----
  const double* xs = point == Point::Average ? xa.get() : point == Point::Pdhg ? xsum.get() : x.get();
  const double* ys = point == Point::Average ? ya.get() : point == Point::Pdhg ? ysum.get() : y.get();
  check(gpk_pdhg_unscale_x(xs, dc, lp.l, lp.u, x_out, n, s), where);
  check(gpk_pdhg_mul(ys, dr, y_out, m, s), where);
  check(gpk_pdhg_row_terms(A0, lp.AL, lp.AU, x_out, y_out, partial, res.get(), s), where);
  check(gpk_pdhg_col_terms(AT0, lp.c, lp.l, lp.u, x_out, y_out, z_out, partial, res.get() + 4, s), where);
  h = cx.to_host(res.get(), 8, where);
  const Kkt fin = kkt_from(h.data(), lp.obj_constant);
  rep.primal_obj = fin.pobj;
  rep.dual_obj = fin.dobj;
  rep.primal_res = fin.rel_p;
  rep.dual_res = fin.rel_d;
  rep.gap = fin.rel_gap;
  rep.primal_weight = omega;
  rep.elapsed_seconds = seconds();
  return rep;
}

}  // namespace gpu_presolver::amd
