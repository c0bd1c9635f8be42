// CPU oracle of the LP presolve/postsolve hot path -- TEST INFRASTRUCTURE.
//
// A serial C++17 restatement of the reference's rule semantics
// (PolyU-IOR/GPU-Presolver, cpp/src/presolve/**), written from the source
// files as text.  Only the test suite, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it: as the parity checker and as the
// serial CPU baseline.  The shipped library never links or calls it.
//
// Reductions that the reference performs with floating-point atomics (row
// shifts, objective deltas) have no defined order there; here they are
// summed in one canonical order that the HIP library also uses:
//   * per-row shifts: row entries in CSR (ascending column) order,
//   * objective deltas: ascending column (or selected-row) order.
// On dyadic data every such sum is exact and any order gives the same bits.
//
// Pinned by: the known-answer cases of the reference's own unit tests
// (tests/golden/reference_kats.json: activity, every rule, scheduler cases
// and postsolve replays), a numpy closed form of row activity, scipy
// transposes/submatrices for the transpose and compaction, and LP
// equivalence checked with HiGHS (optimum preserved by presolve, postsolve
// returning a feasible optimal point).  Parity unpinned beyond this
// restatement (no reference test fixes them): the tiered schedule's record
// on large random LPs, the doubleton batch matching with many overlapping
// candidates, the duplicate-row representative with three or more parallel
// rows, the implied-bound relaxation on non-integral values, and objective
// bits on non-dyadic data.
#pragma once

#include <cstdint>
#include <limits>
#include <string>
#include <vector>

#include "gp_presolve_capi.h"

namespace orc {

constexpr double INF = std::numeric_limits<double>::infinity();

enum RecordType : int32_t {
  FixedCol = 0, FixedColInf = 1, SubCol = 2, DuplicateColumn = 3, DuplicateRow = 4,
  DeletedRow = 5, AddedRow = 6, AddedRows = 7, LhsChange = 8, RhsChange = 9,
  EqToIneq = 10, BoundChangeNoRow = 11, BoundChangeTheRow = 12,
  DoubletonEquation = 13, FmeCol = 14,
};
enum DualMode : uint8_t { DualNone = 0, DualExact = 1, DualMinimal = 2 };

struct Tape {
  std::vector<int32_t> types;
  std::vector<int32_t> index_starts{0};
  std::vector<int32_t> value_starts{0};
  std::vector<uint8_t> dual_modes;
  std::vector<int32_t> indices;
  std::vector<double> vals;
  void append(int32_t type, const std::vector<int32_t>& idx, const std::vector<double>& v,
              uint8_t mode) {
    types.push_back(type);
    indices.insert(indices.end(), idx.begin(), idx.end());
    vals.insert(vals.end(), v.begin(), v.end());
    index_starts.push_back(static_cast<int32_t>(indices.size()));
    value_starts.push_back(static_cast<int32_t>(vals.size()));
    dual_modes.push_back(mode);
  }
};

// Read-only CSR view (host memory).
struct CsrV {
  int32_t rows = 0, cols = 0, nnz = 0;
  const int32_t* ptr = nullptr;
  const int32_t* idx = nullptr;
  const double* val = nullptr;
};

// Owning CSR.
struct Csr {
  int32_t rows = 0, cols = 0;
  std::vector<int32_t> ptr{0};
  std::vector<int32_t> idx;
  std::vector<double> val;
  int32_t nnz() const { return static_cast<int32_t>(idx.size()); }
  CsrV view() const {
    return CsrV{rows, cols, nnz(), ptr.data(), idx.data(), val.data()};
  }
};

Csr copy_csr(const CsrV& a);
Csr transpose(const CsrV& a);

struct LpV {
  CsrV A, AT;
  const double* c = nullptr;
  const double* AL = nullptr;
  const double* AU = nullptr;
  const double* l = nullptr;
  const double* u = nullptr;
  double obj_constant = 0.0;
};

struct Stats {
  int32_t* row_nnz = nullptr;
  uint8_t* singleton_row_mask = nullptr;
  int32_t* singleton_row_col = nullptr;
  double* singleton_row_val = nullptr;
  int32_t* col_nnz = nullptr;
  uint8_t* empty_col_mask = nullptr;
  uint8_t* column_singleton_mask = nullptr;
  int32_t* column_singleton_row = nullptr;
  double* column_singleton_val = nullptr;
};

struct Plan {
  uint8_t* keep_row = nullptr;
  uint8_t* keep_col = nullptr;
  double* c = nullptr;
  double* l = nullptr;
  double* u = nullptr;
  double* AL = nullptr;
  double* AU = nullptr;
  double obj_delta = 0.0;
  bool has_change = false, has_row_action = false, has_col_action = false;
  bool has_infeasible = false, has_unbounded = false;
  bool has_new_A = false;
  Csr new_A;
  Tape tape;
};

using Params = gp_params;
void default_params(Params* p);

// ---- primitives ------------------------------------------------------------
void row_activity(const CsrV& A, const double* l, const double* u, double ztol,
                  double* min_fin, double* max_fin, int32_t* neg_inf, int32_t* pos_inf);
void col_max_abs(const CsrV& AT, double* out);
Csr compact(const CsrV& A, const uint8_t* keep_row, const uint8_t* keep_col);
void sort_rows_in_place(const CsrV& A);
double fixed_col_tape(Tape* tape, const uint8_t* fixed_mask, const uint8_t* excluded,
                      const double* fixed_val, const uint8_t* keep_row, const double* c,
                      const CsrV& AT, int32_t n);

// ---- rules (reference rules/*.cu) -----------------------------------------
void rule_empty_rows(Plan&, const LpV&, const Stats&, const Params&);
void rule_singleton_rows(Plan&, const LpV&, const Stats&, const Params&);
void rule_infeasible_redundant_rows(Plan&, const LpV&, const Stats&, const Params&);
void rule_implied_variable_bounds(Plan&, const LpV&, const Stats&, const Params&);
void rule_duplicate_rows(Plan&, const LpV&, const Stats&, const Params&);
void rule_infeasible_fixed_variables(Plan&, const LpV&, const Params&);
void rule_empty_cols(Plan&, const LpV&, const Stats&, const Params&);
void rule_column_singletons(Plan&, const LpV&, const Stats&, const Params&, bool eq_mode,
                            bool skip_precheck);
void rule_doubleton_equations(Plan&, const LpV&, const Stats&, const Params&);
void rule_dual_fix(Plan&, const LpV&, const Stats&, const Params&, const uint8_t* candidates);
void rule_duplicate_columns(Plan&, const LpV&, const Stats&, const Params&);
void rule_redundant_bounds(Plan&, const LpV&, const Stats&, const Params&);

// ---- presolve / postsolve ----------------------------------------------------
struct Record {
  int32_t m0 = 0, n0 = 0, m1 = 0, n1 = 0;
  std::vector<int32_t> row_red2org, col_red2org;
  std::vector<int32_t> fixed_idx;
  std::vector<double> fixed_val;
  double obj_constant_old = 0.0, obj_constant_new = 0.0;
  Tape tape;
};

struct OwnedLp {
  Csr A, AT;
  std::vector<double> c, AL, AU, l, u;
  double obj_constant = 0.0;
  LpV view() const;
};

struct Result {
  int32_t original_rows = 0, original_cols = 0;
  int32_t reduced_rows = 0, reduced_cols = 0, reduced_nnz = 0;
  int32_t iterations = 0;
  double elapsed_seconds = 0.0;
  double obj_constant_delta = 0.0;
  bool has_infeasible = false, has_unbounded = false;
  Record record;
  OwnedLp reduced;
  bool has_reduced = false;
};

Result presolve(const LpV& lp, const Params& params, bool keep_reduced);
bool has_good_nnz_progress(int32_t before, int32_t after, double ratio);
// Working-LP nnz from which the fast phase starts with the batched singleton
// closure (gpu_presolve.cu:2506: 10,000,000).
void set_batched_closure_min_nnz(int32_t nnz);
int32_t batched_closure_min_nnz();
int64_t batched_closure_runs();
// The duplicate-row hash of every row (the rule's first step), as the
// reference computes it for this matrix (rules/rule_duplicate_rows.cu:78-178,
// 484-497).
void duplicate_row_hashes(const CsrV& A, const uint8_t* keep_row, double ctol, uint64_t* h);

struct RecordV {
  int32_t m0, n0, m1, n1;
  const int32_t* row_red2org;
  const int32_t* col_red2org;
  int32_t fixed_count;
  const int32_t* fixed_idx;
  const double* fixed_val;
  int32_t tape_count;
  const int32_t* types;
  const int32_t* index_starts;
  const int32_t* value_starts;
  const int32_t* indices;
  const double* vals;
};

// x_org, z_org: n0; y_org: m0.  original may be null.
void postsolve(const RecordV& rec, const double* x_red, const double* y_red, const double* z_red,
               const LpV* original, double* x_org, double* y_org, double* z_org);

}  // namespace orc
