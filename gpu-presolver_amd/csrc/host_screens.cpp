// Fail-closed screens of the six structure-specific rules, run where the
// reference's tiered scheduler runs those rules (gpu_presolve.cu:3082-3330):
// the five model-wide analyzers at the start of the bootstrap, and the
// structural L1 prefix screen before and (when the bootstrap changed the LP)
// after the generic bootstrap rules.  Each screen evaluates, on the device,
// the test the reference uses to decide whether its rule acts (or an upper
// bound of it), and throws UnsupportedStructure when the rule would act.
#include <algorithm>
#include <cmath>
#include <string>

#include "gp_host.hpp"
#include "gp_kernels.h"

namespace gpu_presolver::amd {

using presolve::LPInfoGpu;
using presolve::PresolveParams;

namespace {

[[noreturn]] void refuse(const char* rule, const char* flag, const std::string& why) {
  throw UnsupportedStructure(std::string("presolve: the structure-specific rule ") + rule +
                             " would act on this LP (" + why +
                             "); this library does not implement it: set enable_structure_specific_rules or " +
                             flag + " to false");
}

bool well_formed(const LPInfoGpu& lp) {
  return lp.A.rows > 0 && lp.A.cols > 0 && lp.A.nnz > 0 && lp.AT.rows == lp.A.cols && lp.AT.cols == lp.A.rows &&
         lp.AT.nnz == lp.A.nnz;
}

bool enough(int32_t count, int32_t n, int32_t min_count, double min_ratio) {
  return count >= std::max(1, min_count) && static_cast<double>(count) / static_cast<double>(n) + 1.0e-15 >= min_ratio;
}

// rule_covering_cost_dominance.cu:199-326: size gate, full validation of the
// unit covering form, then enough dominated columns.
void screen_covering(const LPInfoGpu& lp, const PresolveParams& prm) {
  const int32_t n = lp.A.cols;
  if (!well_formed(lp) || n < std::max(1, prm.covering_cost_dominance_min_fixed_cols)) return;
  Ctx& c = ctx();
  int32_t* f = c.flags();
  check(gpk_screen_covering(lp.A.rowPtr, lp.A.colVal, lp.A.nzVal, lp.AT.rowPtr, lp.AT.nzVal, lp.c, lp.AL, lp.AU,
                            lp.l, lp.u, lp.A.rows, n, lp.A.nnz, f, c.stream()),
        "covering screen");
  const int32_t* h = c.read_flags(2);
  if (h[0] == 0 && enough(h[1], n, prm.covering_cost_dominance_min_fixed_cols,
                          prm.covering_cost_dominance_min_candidate_ratio))
    refuse("covering cost dominance", "enable_covering_cost_dominance",
           std::to_string(h[1]) + " dominated columns of a unit set-covering LP");
}

// rule_linf_components.cu:45-102: the objective ends in a short positive
// suffix (the tail columns) after a zero, and the columns before it pair up.
bool linf_layout(const LPInfoGpu& lp, const PresolveParams& prm, int32_t* pairs_out) {
  const int32_t m = lp.A.rows, n = lp.A.cols;
  if (n <= 2) return false;
  const int64_t min_pairs = std::max(1, prm.linf_min_pairs);
  if (n < 2 * min_pairs + 1 || m < 2 * min_pairs || lp.A.nnz < 7 * min_pairs) return false;
  const int32_t max_tail = std::max(1, std::min(prm.linf_max_tail_cols, 8));
  const int32_t copy = std::min(n, max_tail + 1);
  const std::vector<double> suffix = ctx().to_host(lp.c + (n - copy), copy, "Linf objective suffix");
  const double last = suffix[static_cast<size_t>(copy - 1)];
  if (!std::isfinite(last) || last <= 0.0) return false;
  int32_t tail = 0;
  for (int32_t k = copy - 1; k >= 0; --k) {
    const double v = suffix[static_cast<size_t>(k)];
    if (!(std::isfinite(v) && v > 0.0 && tail < max_tail)) break;
    ++tail;
  }
  if (tail <= 0 || tail > max_tail || tail >= n || tail == copy) return false;
  const int32_t prefix = n - tail;
  if (suffix[static_cast<size_t>(copy - tail - 1)] != 0.0 || (prefix & 1) != 0) return false;
  if (prefix / 2 < min_pairs) return false;
  *pairs_out = prefix / 2;
  return true;
}

// The sampled pair and merge-row probes of the antipodal (kind 0) and Linf
// (kind 1) detectors (rule_antipodal_components.cu:679-727,
// rule_linf_components.cu:849-914): every sampled pair mirrors and at least 5%
// of the sampled rows are merge rows.
bool pair_probe_accepts(const LPInfoGpu& lp, int32_t kind, int32_t pairs, int32_t probe_pairs) {
  constexpr int32_t kMaxProbeRows = 1024;
  const int32_t probes = std::max(1, std::min(pairs, probe_pairs));
  const int32_t row_probes = std::min(lp.A.rows, kMaxProbeRows);
  Ctx& c = ctx();
  int32_t* f = c.flags();
  check(gpk_screen_pair_probe(kind, lp.A.rowPtr, lp.A.colVal, lp.A.nzVal, lp.AT.rowPtr, lp.AT.colVal, lp.AT.nzVal,
                              lp.c, lp.AL, lp.AU, lp.l, lp.u, lp.A.rows, lp.A.cols, pairs, probes, row_probes, f,
                              c.stream()),
        "pair probe screen");
  const int32_t* h = c.read_flags(2);
  return h[0] == 0 && static_cast<int64_t>(h[1]) * 100 >= static_cast<int64_t>(row_probes) * 5;
}

void screen_linf(const LPInfoGpu& lp, const PresolveParams& prm) {
  int32_t pairs = 0;
  if (!well_formed(lp) || !linf_layout(lp, prm, &pairs)) return;
  if (pair_probe_accepts(lp, 1, pairs, prm.linf_probe_pairs))
    refuse("Linf components", "enable_linf_components",
           "its pair and edge probes accept " + std::to_string(pairs) + " column pairs");
}

void screen_antipodal(const LPInfoGpu& lp, const PresolveParams& prm) {
  const int32_t n = lp.A.cols;
  if (!well_formed(lp) || (n & 1) != 0 || n / 2 < std::max(1, prm.antipodal_min_pairs)) return;
  if (pair_probe_accepts(lp, 0, n / 2, prm.antipodal_probe_pairs))
    refuse("antipodal components", "enable_antipodal_components",
           "its pair and merge-row probes accept " + std::to_string(n / 2) + " column pairs");
}

// rule_bounded_two_row_projection.cu:470-549 (the count is an upper bound of
// the rule's: see k_screens.hip).
void screen_bounded_two_row(const LPInfoGpu& lp, const PresolveParams& prm) {
  if (!well_formed(lp)) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  const int32_t max_row = std::max(2, std::min(8, prm.bounded_two_row_projection_max_row_nnz));
  const double tol = std::max(prm.bounded_two_row_projection_safety_tol,
                              std::max(10.0 * prm.feasibility_tol, 100.0 * prm.zero_tol));
  Ctx& c = ctx();
  DBuf<uint8_t> cand(n, "bounded projection screen candidates");
  DBuf<int32_t> hits(m, "bounded projection screen row hits");
  CopyBatch().zero(hits.get(), m).run("bounded projection screen clear");
  int32_t* f = c.flags();
  check(gpk_screen_bounded_two_row(lp.A.rowPtr, lp.A.colVal, lp.A.nzVal, lp.AT.rowPtr, lp.AT.colVal, lp.c, lp.AL,
                                   lp.AU, lp.l, lp.u, n, max_row, prm.zero_tol, tol, cand, hits, f, c.stream()),
        "bounded projection screen");
  const int32_t selected = c.read_flags(1)[0];
  if (enough(selected, n, prm.bounded_two_row_projection_min_candidates,
             prm.bounded_two_row_projection_min_candidate_ratio))
    refuse("bounded two-row projection", "enable_bounded_two_row_projection",
           "up to " + std::to_string(selected) + " projectable columns");
}

// rule_orphan_mccormick_projection.cu:384-485.
void screen_mccormick(const LPInfoGpu& lp, const PresolveParams& prm) {
  if (!well_formed(lp)) return;
  const int32_t m = lp.A.rows, n = lp.A.cols;
  Ctx& c = ctx();
  DBuf<uint8_t> cand(n, "McCormick screen candidates");
  DBuf<int32_t> ends(2 * static_cast<int64_t>(n), "McCormick screen endpoints");
  DBuf<int32_t> hits(m, "McCormick screen row hits");
  CopyBatch().zero(hits.get(), m).run("McCormick screen clear");
  int32_t* f = c.flags();
  check(gpk_screen_mccormick(lp.A.rowPtr, lp.A.colVal, lp.A.nzVal, lp.AT.rowPtr, lp.AT.colVal, lp.c, lp.AL, lp.AU,
                             lp.l, lp.u, n, cand, ends, hits, f, c.stream()),
        "McCormick screen");
  const int32_t selected = c.read_flags(1)[0];
  if (enough(selected, n, prm.orphan_mccormick_projection_min_candidates,
             prm.orphan_mccormick_projection_min_candidate_ratio))
    refuse("orphan McCormick projection", "enable_orphan_mccormick_projection",
           std::to_string(selected) + " orphan McCormick columns");
}

}  // namespace

void screen_structural_l1(const LPInfoGpu& lp, const PresolveParams& prm) {
  if (!prm.enable_structure_specific_rules || !prm.enable_structural_l1_substitution) return;
  if (lp.A.rows <= 0 || lp.A.cols <= 0) return;
  Ctx& c = ctx();
  int32_t* f = c.flags();
  check(gpk_screen_structural_prefix(lp.A.rowPtr, lp.A.colVal, lp.A.nzVal, lp.AL, lp.AU, lp.A.rows, f, c.stream()),
        "structural L1 screen");
  if (c.read_flags(1)[0] != 0)
    refuse("structural L1 substitution", "enable_structural_l1_substitution",
           "its first rows open an L1 block pattern");
}

void screen_bootstrap_structures(const LPInfoGpu& lp, const PresolveParams& prm) {
  if (!prm.enable_structure_specific_rules) return;
  if (prm.enable_covering_cost_dominance) screen_covering(lp, prm);
  if (prm.enable_linf_components) screen_linf(lp, prm);
  if (prm.enable_antipodal_components) screen_antipodal(lp, prm);
  if (prm.enable_bounded_two_row_projection) screen_bounded_two_row(lp, prm);
  if (prm.enable_orphan_mccormick_projection) screen_mccormick(lp, prm);
  screen_structural_l1(lp, prm);
}

}  // namespace gpu_presolver::amd
