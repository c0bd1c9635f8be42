// Record maintenance (device-resident red2org maps, fixed-column records,
// tape index globalization) and postsolve kernels: scatter of the reduced
// solution, fixed-value restore, reverse tape replay and the column-dual
// projection.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>

#include "gp_device.hpp"
#include "gp_kernels.h"
#include "gp_record_kernels.h"

using namespace gpdev;

namespace {

// ------------------------------------------------------------ record ----
// Removed columns whose bounds are finite and within 1e-9 (gpu_presolve.cu:1005-1043).
__global__ void k_removed_fixed_mask(int32_t n, const uint8_t* keep_col, const double* l, const double* u,
                                     uint8_t* mask) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double lo = l[j], hi = u[j];
  mask[j] = (!keep_col[j] && isfinite(lo) && isfinite(hi) && fabs(lo - hi) <= 1.0e-9) ? 1 : 0;
}

__global__ void k_pack_fixed_records(int32_t n, const uint8_t* mask, const int32_t* prefix, const int32_t* red2org,
                                     const double* l, const double* u, int32_t* out_idx, double* out_val) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || !mask[j]) return;
  const int32_t r = prefix[j];
  out_idx[r] = red2org[j];
  out_val[r] = 0.5 * (l[j] + u[j]);
}

// kind: 1 = row index, 2 = column index, 0 = not an index.  Out-of-range
// values are left unchanged (gpu_presolve.cu:1347-1361).
__global__ void k_map_tape_indices(int32_t count, int32_t* idx, const uint8_t* kind, const int32_t* row_red2org,
                                   int32_t m, const int32_t* col_red2org, int32_t n) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int32_t v = idx[t];
  if (kind[t] == 1) {
    if (v >= 0 && v < m) idx[t] = row_red2org[v];
  } else if (kind[t] == 2) {
    if (v >= 0 && v < n) idx[t] = col_red2org[v];
  }
}

// --------------------------------------------------------- postsolve ----
__global__ void k_scatter(const double* src, const int32_t* red2org, int32_t n_red, double* dst) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_red) dst[red2org[i]] = src[i];
}

__global__ void k_mark(const int32_t* red2org, int32_t n_red, uint8_t* mask) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_red) mask[red2org[i]] = 1;
}

__global__ void k_restore_fixed(const int32_t* idx, const double* val, int32_t count, double* x) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) x[idx[i]] = val[i];
}

__device__ inline bool pfin(double x) { return x > -INFINITY && x < INFINITY; }
__device__ inline bool at_lower(double x, double l, double tol) { return pfin(l) && x <= l + tol; }
__device__ inline bool at_upper(double x, double u, double tol) { return pfin(u) && x >= u - tol; }
__device__ inline double clamp_inf(double x, double lo, double hi) {
  double y = x;
  if (pfin(lo)) y = fmax(y, lo);
  if (pfin(hi)) y = fmin(y, hi);
  return y;
}

// gpu_postsolve.cu:189-210
__device__ double merged_source_value(double lo, double hi, double tol) {
  if (isfinite(lo) && isfinite(hi)) return 0.5 * (lo + hi);
  if (isfinite(lo)) {
    if (0.0 > lo + tol) return 0.0;
    return lo + fmax(1.0, fabs(lo) * fmax(tol, 1.0e-12));
  }
  if (isfinite(hi)) {
    if (0.0 < hi - tol) return 0.0;
    return hi - fmax(1.0, fabs(hi) * fmax(tol, 1.0e-12));
  }
  return 0.0;
}

// gpu_postsolve.cu:240-282
__device__ void recover_dup_col(double y, double z_to, double r, double sl, double su, double tl, double tu,
                                double tol, double* x_from, double* x_to) {
  double lo = sl, hi = su;
  if (isfinite(tu)) { const double b = (y - tu) / r; if (r > 0.0) lo = fmax(lo, b); else hi = fmin(hi, b); }
  if (isfinite(tl)) { const double b = (y - tl) / r; if (r > 0.0) hi = fmin(hi, b); else lo = fmax(lo, b); }
  double lower = -INFINITY, upper = INFINITY;
  if (r > 0.0) {
    if (isfinite(tl) && isfinite(sl)) lower = tl + r * sl;
    if (isfinite(tu) && isfinite(su)) upper = tu + r * su;
  } else {
    if (isfinite(tl) && isfinite(su)) lower = tl + r * su;
    if (isfinite(tu) && isfinite(sl)) upper = tu + r * sl;
  }
  if (z_to > tol && isfinite(lower) && fabs(y - lower) <= tol) *x_from = r > 0.0 ? sl : su;
  else if (z_to < -tol && isfinite(upper) && fabs(y - upper) <= tol) *x_from = r > 0.0 ? su : sl;
  else *x_from = merged_source_value(lo, hi, tol);
  *x_to = y - r * (*x_from);
}

enum : uint8_t { ST_ZERO = 0, ST_NONNEG = 1, ST_NONPOS = 2, ST_FREE = 3 };
__device__ uint8_t row_state(double act, double lo, double hi, double tol) {
  const bool al = pfin(lo) && act <= lo + tol;
  const bool au = pfin(hi) && act >= hi - tol;
  if (al && au) return ST_FREE;
  if (al) return ST_NONNEG;
  if (au) return ST_NONPOS;
  return ST_ZERO;
}
__device__ void state_interval(uint8_t s, double* lo, double* hi) {
  if (s == ST_FREE) { *lo = -INFINITY; *hi = INFINITY; }
  else if (s == ST_NONNEG) { *lo = 0.0; *hi = INFINITY; }
  else if (s == ST_NONPOS) { *lo = -INFINITY; *hi = 0.0; }
  else { *lo = 0.0; *hi = 0.0; }
}
__device__ void deleted_dual_interval(double yp, double r, uint8_t ks, double* lo, double* hi) {
  const double piv = r * yp;
  if (ks == ST_FREE) { *lo = -INFINITY; *hi = INFINITY; }
  else if (ks == ST_ZERO) { *lo = piv; *hi = piv; }
  else if (ks == ST_NONNEG) { if (r > 0.0) { *lo = -INFINITY; *hi = piv; } else { *lo = piv; *hi = INFINITY; } }
  else { if (r > 0.0) { *lo = piv; *hi = INFINITY; } else { *lo = -INFINITY; *hi = piv; } }
}
// gpu_postsolve.cu:409-469
__device__ bool doc_dual_split(double yp, double r, double kal, double kau, double dsl, double dsu, double tol,
                               double* op, double* od) {
  if (fabs(yp) <= tol) { *op = 0.0; *od = 0.0; return true; }
  const double bl = fmax(kal, dsl), bu = fmin(kau, dsu);
  if (!(bl <= bu + tol)) { *op = yp; *od = 0.0; return false; }
  const bool lt = bl > kal + tol, ut = bu < kau - tol;
  if (lt && ut) { *op = yp; *od = 0.0; return false; }
  if (!lt && !ut) { *op = yp; *od = 0.0; return true; }
  if (yp > 0.0) {
    if (lt) { *op = 0.0; *od = r * yp; } else { *op = yp; *od = 0.0; }
    return true;
  }
  if (lt) { *op = yp; *od = 0.0; } else { *op = 0.0; *od = r * yp; }
  return true;
}
// gpu_postsolve.cu:471-515
__device__ void dbl_recover_z(double xk, double zk_hat, double al, double okl, double oku, double nkl, double nku,
                              double ls, double us, double tol, double* zk, double* ze) {
  if (fabs(al) <= tol) { *zk = zk_hat; *ze = 0.0; return; }
  const bool atl = pfin(nkl) && xk <= nkl + tol;
  const bool atu = pfin(nku) && xk >= nku - tol;
  if (!atl && !atu) { *zk = 0.0; *ze = 0.0; return; }
  double src = atl ? ls : us;
  if (src == 0.0) {
    if (atl) src = (pfin(okl) && fabs(nkl - okl) <= tol) ? 1.0 : 2.0;
    else src = (pfin(oku) && fabs(nku - oku) <= tol) ? 1.0 : 2.0;
  }
  if (src == 2.0) { *zk = 0.0; *ze = zk_hat / al; }
  else { *zk = zk_hat; *ze = 0.0; }
}

// s (+/-)= sum over q in [0, count) with idx[q] != skip of val[q] * vec[idx[q]],
// accumulated in q order exactly as the plain loop would (every product
// rounded, then added or subtracted one at a time).  The indices, values and
// gathered entries of eight terms are loaded before they are used, so a long
// column costs count / 8 memory round trips instead of count.
template <bool kSub, int U = 8>
__device__ inline double dot_in_order(double s, const int32_t* __restrict__ idx, const double* __restrict__ val,
                                      const double* __restrict__ vec, int32_t count, int32_t skip) {
  int32_t q = 0;
  for (; q + U <= count; q += U) {
    int32_t j[U];
    double a[U], v[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      j[k] = idx[q + k];
      a[k] = val[q + k];
    }
#pragma unroll
    for (int k = 0; k < U; ++k) v[k] = vec[j[k]];
#pragma unroll
    for (int k = 0; k < U; ++k)
      if (j[k] != skip) s = kSub ? s - a[k] * v[k] : s + a[k] * v[k];
  }
  for (; q < count; ++q) {
    const int32_t j = idx[q];
    if (j != skip) s = kSub ? s - val[q] * vec[j] : s + val[q] * vec[j];
  }
  return s;
}

// Where a record's indices and values sit in the tape (immutable, so it can
// be loaded ahead of the record's turn).
// The header also carries the record's first three index words (-1 past
// its end), so a prefetched header leaves the record's variables one load
// away.
struct RecHead {
  int32_t type, i0, i1, v0, v1;
  int32_t a0, a1, a2;
};
__device__ inline RecHead load_head(const gp_tape_dev& T, int32_t k) {
  RecHead h{T.types[k], T.index_starts[k], T.index_starts[k + 1] - 1, T.value_starts[k], T.value_starts[k + 1] - 1,
            -1, -1, -1};
  if (h.i0 <= h.i1) h.a0 = T.indices[h.i0];
  if (h.i0 + 1 <= h.i1) h.a1 = T.indices[h.i0 + 1];
  if (h.i0 + 2 <= h.i1) h.a2 = T.indices[h.i0 + 2];
  return h;
}

// Dot products inside a record keep this many entries' loads in flight.
constexpr int kRecDotU = 32;

// One record of the reverse replay (_kernel_replay_postsolve_tape,
// gpu_postsolve.cu:605-1046).
__device__ void replay_one(const gp_tape_dev& T, const gp_orig_dev& O, double* x, double* y, double* z,
                           uint8_t* zret, const RecHead& h, bool primal, bool dual, double tol) {
  const bool has_orig = O.present != 0;
  const int32_t type = h.type;
  const int32_t i0 = h.i0, i1 = h.i1;
  const int32_t v0 = h.v0, v1 = h.v1;
  const int32_t* I = T.indices;
  const double* V = T.vals;
  if (type == 3) {  // DuplicateColumn
    const int32_t from = I[i0], to = I[i0 + 1];
    const double r = V[v0];
    if (primal) {
      double xf = 0.0, xt = x[to];
      recover_dup_col(x[to], z ? z[to] : 0.0, r, V[v0 + 1], V[v0 + 2], V[v0 + 3], V[v0 + 4], tol, &xf, &xt);
      x[from] = xf;
      x[to] = xt;
    }
    if (dual && z) { z[from] = r * z[to]; if (zret) zret[from] = 1; }
  } else if (type == 13) {  // DoubletonEquation
    const int32_t ec = h.a0, kc = h.a1, row = h.a2;
    const double ecoef = V[v0], kcoef = V[v0 + 1], rhs = V[v0 + 2], eobj = V[v0 + 9];
    const double al = -kcoef / ecoef, be = rhs / ecoef;
    if (primal) x[ec] = be + al * x[kc];
    if (dual && z && y && has_orig) {
      // The column bounds and c[ec] load together with x[kc] and z[kc].
      const int32_t qb = O.at_ptr[ec], qe = O.at_ptr[ec + 1];
      const double cec = O.c[ec];
      const double ls = v1 >= v0 + 10 ? V[v0 + 10] : 0.0;
      const double us = v1 >= v0 + 11 ? V[v0 + 11] : 0.0;
      double zk = 0.0, ze = 0.0;
      dbl_recover_z(x[kc], z[kc], al, V[v0 + 5], V[v0 + 6], V[v0 + 7], V[v0 + 8], ls, us, tol, &zk, &ze);
      const bool payload = i1 >= i0 + 3 && v1 >= v0 + 11;
      double so = 0.0;
      double yrow;
      if (payload) {
        so = dot_in_order<false, kRecDotU>(so, I + i0 + 4, V + v0 + 12, y, I[i0 + 3], -1);
        yrow = (eobj - so - ze) / ecoef;
      } else {
        so = dot_in_order<false, kRecDotU>(so, O.at_idx + qb, O.at_val + qb, y, qe - qb, row);
        yrow = (cec - so - ze) / ecoef;
      }
      y[row] = yrow;
      z[kc] = zk;
      z[ec] = ze;
      if (zret) { zret[kc] = 1; zret[ec] = 1; }
    } else if (dual) {
      if (z) { z[ec] = 0.0; if (zret) zret[ec] = 1; }
      if (y && row >= 0) y[row] = eobj / ecoef;
    }
  } else if (type == 2) {  // SubCol
    const int32_t ec = h.a0, row = h.a1, sc = h.a2;
    const double piv = V[v0], rhs = V[v0 + 1], eobj = V[v0 + 4];
    const bool row_deleted = V[v0 + 5] != 0.0;
    const double part = dot_in_order<false, kRecDotU>(0.0, I + i0 + 3, V + v0 + 6, x, sc, -1);
    if (primal) x[ec] = (rhs - part) / piv;
    if (dual && y && row >= 0) {
      if (row_deleted) y[row] = eobj / piv; else y[row] += eobj / piv;
    }
    if (dual && z) {
      z[ec] = (row_deleted || !y) ? 0.0 : eobj - piv * y[row];
      if (zret) zret[ec] = 1;
    }
  } else if (type == 0) {  // FixedCol
    const int32_t col = h.a0;
    if (primal) x[col] = V[v0];
    if (dual && z) {
      double zj = V[v0 + 1];
      if (y) zj = dot_in_order<true, kRecDotU>(zj, I + i0 + 1, V + v0 + 2, y, i1 - i0, -1);
      z[col] = zj;
      if (zret) zret[col] = 1;
    }
  } else if (type == 1) {  // FixedColInf
    const bool pos_inf = I[i0] > 0;
    const int32_t col = I[i0 + 1];
    const int32_t nrows = static_cast<int32_t>(llround(V[v0]));
    double ext = V[v0 + 1];
    int32_t ip = i0 + 2, vp = v0 + 2;
    if (primal) {
      for (int32_t r = 0; r < nrows; ++r) {
        const int32_t len = I[ip];
        double side = V[vp], coef = 0.0;
        for (int32_t t = 0; t < len; ++t) {
          const int32_t rc = I[ip + 1 + t];
          const double rv = V[vp + 1 + t];
          if (rc == col) coef = rv; else side -= rv * x[rc];
        }
        const double cand = side / coef;
        ext = pos_inf ? fmax(ext, cand) : fmin(ext, cand);
        ip += len + 1;
        vp += len + 1;
      }
      x[col] = ext;
    }
    if (dual && z) { z[col] = 0.0; if (zret) zret[col] = 1; }
  } else if (type == 14) {  // FmeCol
    const int32_t K = I[i0], ec = I[i0 + 1];
    const int32_t L = static_cast<int32_t>(llround(V[v0]));
    int32_t cc = i0 + 3 + K, vp = v0 + 1;
    double lo = -INFINITY, hi = INFINITY;
    bool hl = false, hh = false;
    if (primal) {
      for (int32_t s = 0; s < L; ++s) {
        const double orient = V[vp];
        const int32_t len = static_cast<int32_t>(llround(V[vp + 1]));
        double cand = V[vp + 2];
        for (int32_t t = 0; t < len; ++t) cand += V[vp + 3 + t] * x[I[cc++]];
        if (orient > 0.0) { lo = fmax(lo, cand); hl = true; }
        else { hi = fmin(hi, cand); hh = true; }
        vp += 3 + len;
      }
      x[ec] = hl ? lo : (hh ? hi : 0.0);
    }
    if (dual && z) { z[ec] = 0.0; if (zret) zret[ec] = 1; }
    if (dual && y) for (int32_t t = 0; t < K; ++t) y[I[i0 + 2 + t]] = 0.0;
  } else if (dual && type == 10) {  // EqToIneq
    if (y) y[I[i0]] += V[v0];
  } else if (dual && type == 4 && i1 >= i0 + 1 && v1 >= v0 + 4) {  // DuplicateRow
    const int32_t kr = I[i0], dr = I[i0 + 1];
    const double r = V[v0], kal = V[v0 + 1], kau = V[v0 + 2], dal = V[v0 + 3], dau = V[v0 + 4];
    if (fabs(r) <= tol || !y) return;
    const double a = dal * r, b = dau * r;
    const double dsl = r >= 0.0 ? a : b, dsu = r >= 0.0 ? b : a;
    if (dsl <= kal + tol && kau <= dsu + tol) { y[dr] = 0.0; return; }
    if (has_orig) {
      const int32_t kb = O.a_ptr[kr], db = O.a_ptr[dr];
      const double ap = dot_in_order<false, kRecDotU>(0.0, O.a_idx + kb, O.a_val + kb, x, O.a_ptr[kr + 1] - kb, -1);
      const double ad = dot_in_order<false, kRecDotU>(0.0, O.a_idx + db, O.a_val + db, x, O.a_ptr[dr + 1] - db, -1);
      const uint8_t ks = row_state(ap, kal, kau, tol), ds = row_state(ad, dal, dau, tol);
      double ydl, ydh, ypl, yph;
      state_interval(ds, &ydl, &ydh);
      deleted_dual_interval(y[kr], r, ks, &ypl, &yph);
      const double sl = fmax(ydl, ypl), sh = fmin(ydh, yph);
      if (sl <= sh + tol) {
        const double ydn = clamp_inf(0.0, sl, sh);
        y[kr] = y[kr] - ydn / r;
        y[dr] = ydn;
        return;
      }
      double ypd = y[kr], ydd = 0.0;
      if (doc_dual_split(y[kr], r, kal, kau, dsl, dsu, tol, &ypd, &ydd)) { y[kr] = ypd; y[dr] = ydd; }
    }
  } else if (dual && type == 5 && has_orig && i1 >= i0 + 1 && v1 >= v0 + 2) {  // DeletedRow
    const int32_t row = h.a0, col = h.a1;
    const double oal = V[v0], oau = V[v0 + 1], aij = V[v0 + 2];
    const int32_t qb = O.at_ptr[col], qe = O.at_ptr[col + 1];
    if (fabs(aij) <= tol) return;
    double gamma = O.c[col];
    if (y) gamma = dot_in_order<true, kRecDotU>(gamma, O.at_idx + qb, O.at_val + qb, y, qe - qb, row);
    const double xj = x[col], zbar = z[col];
    const double act = aij * xj;
    const bool ral = pfin(oal) && act <= oal + tol;
    const bool rau = pfin(oau) && act >= oau - tol;
    double izl = -INFINITY, izh = INFINITY;
    if (!(ral && rau)) {
      if (ral) { if (aij > 0.0) izh = gamma; else izl = gamma; }
      else if (rau) { if (aij > 0.0) izl = gamma; else izh = gamma; }
      else { izl = gamma; izh = gamma; }
    }
    const bool atl = at_lower(xj, O.l[col], tol), atu = at_upper(xj, O.u[col], tol);
    double zl = izl, zh = izh;
    if (!atl && !atu) {
      if ((pfin(izl) && izl > tol) || (pfin(izh) && izh < -tol)) { zl = gamma; zh = gamma; }
      else { zl = 0.0; zh = 0.0; }
    } else if (atl && !atu) {
      zl = fmax(zl, 0.0);
    } else if (!atl && atu) {
      zh = fmin(zh, 0.0);
    }
    if (pfin(zl) && pfin(zh) && zl > zh + tol) return;
    const double zn = clamp_inf(zbar, zl, zh);
    z[col] = zn;
    if (y) y[row] = (gamma - zn) / aij;
    if (zret) zret[col] = 1;
  } else if (dual && type == 12 && has_orig && i1 >= i0 + 1) {  // BoundChangeTheRow
    const int32_t col = h.a0, row = h.a1;
    // Every load that depends only on the record goes out at once; the
    // decisions below then read registers.
    const uint8_t col_set = zret ? zret[col] : 1;
    const double ol = V[v0], ou = V[v0 + 1], nl = V[v0 + 2], nu = V[v0 + 3];
    const double xj = x[col], zj = z[col];
    const int32_t qb = O.a_ptr[row], qe = O.a_ptr[row + 1];
    if (!col_set) return;
    const bool lc = nl > ol + tol, uc = nu < ou - tol;
    if (lc == uc) return;
    const double implied = lc ? nl : nu;
    const double other = lc ? ou : ol;
    const bool other_is_lower = !lc;
    if (pfin(other) && fabs(xj - other) <= tol) {
      const bool ok = other_is_lower ? (zj >= -tol) : (zj <= tol);
      if (ok) return;
    }
    if (fabs(xj - implied) > tol) return;
    // The row's entries are read U at a time with every load in flight
    // together (the columns of a row are distinct, so the updates of one
    // chunk are independent); the results are those of the entry-by-entry
    // loops.  A row of at most U entries is loaded once for both passes.
    constexpr int U = 16;
    if (qe <= qb) return;  // an empty row holds no a_ij
    if (qe - qb <= U) {
      int32_t kc[U];
      double a[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int32_t q = min(qb + k, qe - 1);
        kc[k] = O.a_idx[q];
        a[k] = O.a_val[q];
      }
      double aij = 0.0;
#pragma unroll
      for (int k = U - 1; k >= 0; --k)
        if (qb + k < qe && kc[k] == col) aij = a[k];
      if (fabs(aij) <= tol) return;
      bool on[U];
      double zk[U];
#pragma unroll
      for (int k = 0; k < U; ++k) on[k] = qb + k < qe && kc[k] != col && (!zret || zret[kc[k]]);
#pragma unroll
      for (int k = 0; k < U; ++k) zk[k] = z[kc[k]];
      if (y) y[row] += zj / aij;
#pragma unroll
      for (int k = 0; k < U; ++k)
        if (on[k]) z[kc[k]] = zk[k] - (a[k] / aij) * zj;
      z[col] = 0.0;
      if (zret) zret[col] = 1;
      return;
    }
    double aij = 0.0;
    for (int32_t q0 = qb; q0 < qe; q0 += U) {
      int32_t j[U];
      double a[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int32_t q = min(q0 + k, qe - 1);
        j[k] = O.a_idx[q];
        a[k] = O.a_val[q];
      }
      bool hit = false;
#pragma unroll
      for (int k = U - 1; k >= 0; --k)
        if (q0 + k < qe && j[k] == col) {
          aij = a[k];
          hit = true;
        }
      if (hit) break;
    }
    if (fabs(aij) <= tol) return;
    if (y) y[row] += zj / aij;
    for (int32_t q0 = qb; q0 < qe; q0 += U) {
      int32_t kc[U];
      double a[U], zk[U];
      bool on[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int32_t q = min(q0 + k, qe - 1);
        kc[k] = O.a_idx[q];
        a[k] = O.a_val[q];
      }
#pragma unroll
      for (int k = 0; k < U; ++k) on[k] = q0 + k < qe && kc[k] != col && (!zret || zret[kc[k]]);
#pragma unroll
      for (int k = 0; k < U; ++k) zk[k] = z[kc[k]];
#pragma unroll
      for (int k = 0; k < U; ++k)
        if (on[k]) z[kc[k]] = zk[k] - (a[k] / aij) * zj;
    }
    z[col] = 0.0;
    if (zret) zret[col] = 1;
  }
}

// Records begin..end-1 in reverse, one lane (the reference's order).
__global__ void k_replay(gp_tape_dev T, gp_orig_dev O, double* x, double* y, double* z, uint8_t* zret,
                         int32_t begin, int32_t end, int32_t primal, int32_t dual, double tol) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int32_t k = end - 1; k >= begin; --k)
    replay_one(T, O, x, y, z, zret, load_head(T, k), primal != 0, dual != 0, tol);
}

// One level of a dependency-level schedule: its records touch disjoint
// variables and are replayed one per thread.
__global__ void __launch_bounds__(kBlock) k_replay_list(gp_tape_dev T, gp_orig_dev O, double* x, double* y, double* z, uint8_t* zret,
                              const int32_t* records, int32_t count, int32_t primal, int32_t dual, double tol) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) replay_one(T, O, x, y, z, zret, load_head(T, records[i]), primal != 0, dual != 0, tol);
}

// A run of small levels in one workgroup: the threads share each level's
// records and meet at a barrier before the next level.  All waves of the
// workgroup run on one CU, so the barrier's workgroup-scope fence makes one
// level's global stores visible to the next level's loads.
// The level bounds of up to kLevelChunk levels sit in LDS, and each thread
// loads the header of its record of the next level while the current level
// runs, so a level's critical path starts at the tape's indices and values.
constexpr int kLevelBlock = 512;
constexpr int kLevelChunk = 4096;
__global__ void __launch_bounds__(kLevelBlock)
    k_replay_levels(gp_tape_dev T, gp_orig_dev O, double* x, double* y, double* z, uint8_t* zret,
                    const int32_t* records, const int32_t* level_start, int32_t nlevels, int32_t primal,
                    int32_t dual, double tol) {
  __shared__ int32_t s_start[kLevelChunk + 2];
  const int32_t tid = static_cast<int32_t>(threadIdx.x);
  for (int32_t c0 = 0; c0 < nlevels; c0 += kLevelChunk) {
    const int32_t cl = min(kLevelChunk, nlevels - c0);
    // level_start[c0 .. c0 + cl + 1]: the bounds of this chunk's levels and of the next level
    const int32_t have = min(cl + 2, nlevels + 1 - c0);
    for (int32_t t = tid; t < have; t += kLevelBlock) s_start[t] = level_start[c0 + t];
    __syncthreads();
    int32_t b = s_start[0], e = s_start[1];
    RecHead cur{};
    if (b + tid < e) cur = load_head(T, records[b + tid]);
    for (int32_t l = 0; l < cl; ++l) {
      const bool more = l + 1 < have - 1;  // the next level's bounds are in LDS
      const int32_t nb = more ? s_start[l + 1] : 0, ne = more ? s_start[l + 2] : 0;
      RecHead nxt{};
      if (nb + tid < ne) nxt = load_head(T, records[nb + tid]);
      if (b + tid < e) replay_one(T, O, x, y, z, zret, cur, primal != 0, dual != 0, tol);
      for (int32_t i = b + tid + kLevelBlock; i < e; i += kLevelBlock)
        replay_one(T, O, x, y, z, zret, load_head(T, records[i]), primal != 0, dual != 0, tol);
      __syncthreads();
      b = nb;
      e = ne;
      cur = nxt;
    }
    __syncthreads();  // s_start is reloaded for the next chunk
  }
}

// gpu_postsolve.cu:162-180, 568-603: z_j from r = c_j - (A^T y)_j and the
// bound state of x_j.  (A^T y)_j is summed in AT order; columns longer than
// kLongCol are left to k_project_duals_long, which loads a column 64 entries
// at a time across a wave and adds the products in the same order.
constexpr int32_t kLongCol = 64;

__device__ inline double project_dual(const gp_orig_dev& O, int32_t col, double aty, const double* x, double tol) {
  const double r = O.c[col] - aty;
  const double xj = x[col], lj = O.l[col], uj = O.u[col];
  const bool atl = at_lower(xj, lj, tol), atu = at_upper(xj, uj, tol);
  double zj;
  if (atl && atu) zj = r;
  else if (!atl && !atu) zj = 0.0;
  else if (atl) zj = fmax(r, 0.0);
  else zj = fmin(r, 0.0);
  const double eps = fmin(tol, 1.0e-12);
  if (zj > eps && !isfinite(lj)) zj = 0.0;
  else if (zj < -eps && !isfinite(uj)) zj = 0.0;
  return zj;
}

__global__ void k_project_duals(gp_orig_dev O, const uint8_t* prot, const double* x, const double* y, double* z,
                                int32_t n, double tol, int32_t* long_count, int32_t* long_cols) {
  const int32_t col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n || (prot && prot[col])) return;
  const int32_t b = O.at_ptr[col], e = O.at_ptr[col + 1];
  if (e - b > kLongCol) {
    long_cols[atomicAdd(long_count, 1)] = col;
    return;
  }
  const double aty = dot_in_order<false>(0.0, O.at_idx + b, O.at_val + b, y, e - b, -1);
  z[col] = project_dual(O, col, aty, x, tol);
}

// Lane t's double, in every lane (t the same in all lanes).
__device__ inline double lane_f64(double v, int t) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), t);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), t);
  return __hiloint2double(hi, lo);
}

// One wave per listed column; a fixed grid whose waves stride over the list.
// The column's products are formed kLongSlices x 64 at a time (all loads of
// the batch in flight together) and added to the running sum one by one in
// AT order, each read from its lane into scalar registers, so the serial
// chain costs one add per entry.
constexpr int kLongSlices = 4;
constexpr int kLongGrid = 2048;
__global__ void k_project_duals_long(gp_orig_dev O, const double* x, const double* y, double* z, double tol,
                                     const int32_t* long_count, const int32_t* long_cols) {
  const int lane = lane_id();
  const int32_t waves = gridDim.x * (blockDim.x / 64);
  const int32_t count = *long_count;
  for (int32_t w = blockIdx.x * (blockDim.x / 64) + static_cast<int32_t>(threadIdx.x / 64); w < count; w += waves) {
    const int32_t col = long_cols[w];
    const int32_t b = O.at_ptr[col], e = O.at_ptr[col + 1];
    double aty = 0.0;
    for (int32_t base = b; base < e; base += 64 * kLongSlices) {
      int32_t j[kLongSlices];
      double a[kLongSlices], p[kLongSlices];
#pragma unroll
      for (int u = 0; u < kLongSlices; ++u) {
        const int32_t q = base + 64 * u + lane;
        j[u] = q < e ? O.at_idx[q] : -1;
        a[u] = q < e ? O.at_val[q] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kLongSlices; ++u) p[u] = j[u] >= 0 ? a[u] * y[j[u]] : 0.0;
#pragma unroll
      for (int u = 0; u < kLongSlices; ++u) {
        const int32_t s0 = base + 64 * u;
        if (s0 + 64 <= e) {
#pragma unroll
          for (int t = 0; t < 64; ++t) aty += lane_f64(p[u], t);
        } else {
          for (int t = 0; t < e - s0; ++t) aty += lane_f64(p[u], t);
        }
      }
    }
    if (lane == 0) z[col] = project_dual(O, col, aty, x, tol);
  }
}

// Lengths and packed copies of selected CSR lines (the original rows and
// columns that the host schedules the dual replay by).
__global__ void k_seg_lengths(const int32_t* ptr, const int32_t* list, int32_t count, int32_t* len) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) len[i] = ptr[list[i] + 1] - ptr[list[i]];
}

__global__ void k_seg_copy(const int32_t* ptr, const int32_t* idx, const int32_t* list, int32_t count,
                           const int32_t* off, int32_t* out) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int32_t b = ptr[list[i]], e = ptr[list[i] + 1];
  int32_t o = off[i];
  for (int32_t q = b; q < e; ++q) out[o++] = idx[q];
}

// Columns whose dual the replay determines exactly (gpu_postsolve.cu:1370-1402).
__global__ void k_protect(gp_tape_dev T, int32_t n, uint8_t* prot) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= T.count) return;
  const int32_t i0 = T.index_starts[k];
  const int32_t t = T.types[k];
  int32_t a = -1, b = -1;
  if (t == 2) a = T.indices[i0];
  else if (t == 13) { a = T.indices[i0]; b = T.indices[i0 + 1]; }
  else if (t == 1) a = T.indices[i0 + 1];
  else if (t == 3) a = T.indices[i0];
  if (a >= 0 && a < n) prot[a] = 1;
  if (b >= 0 && b < n) prot[b] = 1;
}

}  // namespace

#define GPK_LAUNCH(kernel, n, ...)                                                      \
  do {                                                                                  \
    if ((n) > 0) hipLaunchKernelGGL(kernel, dim3(grid_for(n)), dim3(kBlock), 0, stream, __VA_ARGS__); \
    return hipGetLastError();                                                           \
  } while (0)

extern "C" {

hipError_t gpk_removed_fixed_mask(int32_t n, const uint8_t* keep_col, const double* l, const double* u,
                                  uint8_t* mask, hipStream_t stream) {
  GPK_LAUNCH(k_removed_fixed_mask, n, n, keep_col, l, u, mask);
}
hipError_t gpk_pack_fixed_records(int32_t n, const uint8_t* mask, const int32_t* prefix, const int32_t* red2org,
                                  const double* l, const double* u, int32_t* out_idx, double* out_val,
                                  hipStream_t stream) {
  GPK_LAUNCH(k_pack_fixed_records, n, n, mask, prefix, red2org, l, u, out_idx, out_val);
}
hipError_t gpk_map_tape_indices(int32_t count, int32_t* idx, const uint8_t* kind, const int32_t* row_red2org,
                                int32_t m, const int32_t* col_red2org, int32_t n, hipStream_t stream) {
  GPK_LAUNCH(k_map_tape_indices, count, count, idx, kind, row_red2org, m, col_red2org, n);
}
hipError_t gpk_scatter(const double* src, const int32_t* red2org, int32_t n_red, double* dst, hipStream_t stream) {
  GPK_LAUNCH(k_scatter, n_red, src, red2org, n_red, dst);
}
hipError_t gpk_mark(const int32_t* red2org, int32_t n_red, uint8_t* mask, hipStream_t stream) {
  GPK_LAUNCH(k_mark, n_red, red2org, n_red, mask);
}
hipError_t gpk_restore_fixed(const int32_t* idx, const double* val, int32_t count, double* x, hipStream_t stream) {
  GPK_LAUNCH(k_restore_fixed, count, idx, val, count, x);
}
hipError_t gpk_replay(gp_tape_dev tape, gp_orig_dev orig, double* x, double* y, double* z, uint8_t* zret,
                      int32_t begin, int32_t end, int32_t primal, int32_t dual, double tol, hipStream_t stream) {
  if (end <= begin) return hipSuccess;
  hipLaunchKernelGGL(k_replay, dim3(1), dim3(64), 0, stream, tape, orig, x, y, z, zret, begin, end, primal, dual, tol);
  return hipGetLastError();
}
hipError_t gpk_replay_list(gp_tape_dev tape, gp_orig_dev orig, double* x, double* y, double* z, uint8_t* zret,
                           const int32_t* records, int32_t count, int32_t primal, int32_t dual, double tol,
                           hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_replay_list, dim3(grid_for(count)), dim3(kBlock), 0, stream, tape, orig, x, y, z, zret, records,
                     count, primal, dual, tol);
  return hipGetLastError();
}
hipError_t gpk_replay_levels(gp_tape_dev tape, gp_orig_dev orig, double* x, double* y, double* z, uint8_t* zret,
                             const int32_t* records, const int32_t* level_start, int32_t nlevels, int32_t primal,
                             int32_t dual, double tol, hipStream_t stream) {
  if (nlevels <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_replay_levels, dim3(1), dim3(kLevelBlock), 0, stream, tape, orig, x, y, z, zret, records,
                     level_start, nlevels, primal, dual, tol);
  return hipGetLastError();
}
hipError_t gpk_seg_lengths(const int32_t* ptr, const int32_t* list, int32_t count, int32_t* len, hipStream_t stream) {
  GPK_LAUNCH(k_seg_lengths, count, ptr, list, count, len);
}
hipError_t gpk_seg_copy(const int32_t* ptr, const int32_t* idx, const int32_t* list, int32_t count,
                        const int32_t* off, int32_t* out, hipStream_t stream) {
  GPK_LAUNCH(k_seg_copy, count, ptr, idx, list, count, off, out);
}
hipError_t gpk_protect_columns(gp_tape_dev tape, int32_t n, uint8_t* prot, hipStream_t stream) {
  GPK_LAUNCH(k_protect, tape.count, tape, n, prot);
}
hipError_t gpk_project_duals(gp_orig_dev orig, const uint8_t* prot, const double* x, const double* y, double* z,
                             int32_t n, double tol, int32_t* long_count, int32_t* long_cols, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_project_duals, dim3(grid_for(n)), dim3(kBlock), 0, stream, orig, prot, x, y, z, n, tol,
                     long_count, long_cols);
  hipLaunchKernelGGL(k_project_duals_long, dim3(kLongGrid), dim3(kBlock), 0, stream, orig, x, y, z, tol, long_count,
                     long_cols);
  return hipGetLastError();
}

}  // extern "C"
