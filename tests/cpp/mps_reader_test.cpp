// The reader's C++ interface as a caller sees it (include/mps_reader/),
// with the known answers of the reference reader's own tests transcribed:
// cpp/mps_reader/tests/test_sparse_matrix.cpp:5-50,
// test_lp_model.cpp:6-41 and test_mpsreader.cpp:10-235.  Host only.  Prints
// "ok <case>" per passing case; exits 1 at the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mps_reader/mpsreader.hpp"

using mps_reader::CscMatrix;
using mps_reader::CsrMatrix;
using mps_reader::LpModel;
using mps_reader::read_mps;

namespace {

const double kInf = std::numeric_limits<double>::infinity();

void require(bool ok, const char* what) {
  if (!ok) throw std::runtime_error(what);
}

template <class F>
bool throws(F f) {
  try {
    f();
  } catch (const std::exception&) {
    return true;
  }
  return false;
}

mps_reader::MpsModel parse(const std::string& text) {
  std::istringstream in(text);
  return read_mps(in);
}

std::string numeric_model(const std::string& columns = "    X OBJ 1 R 1\n", const std::string& rhs = "    RHS1 R 1\n",
                          const std::string& bounds = " UP BND1 X 2\n") {
  return "NAME validation\nROWS\n N OBJ\n L R\nCOLUMNS\n" + columns + "RHS\n" + rhs + "BOUNDS\n" + bounds + "ENDATA\n";
}

bool rejected(const std::string& text) {
  return throws([&] { parse(text); });
}

void csr_rejects_bad_row_pointer_size() {
  require(throws([] { CsrMatrix m(2, 3, {0, 1}, {0}, {1.0}); }), "bad row pointer size accepted");
}

void csc_to_csr_preserves_entries() {
  const CsrMatrix csr = CscMatrix(3, 4, {0, 2, 3, 3, 5}, {0, 2, 1, 0, 2}, {1.0, 5.0, 2.0, 3.0, 7.0}).to_csr();
  require(csr.rows() == 3 && csr.cols() == 4, "shape");
  require(csr.row_ptr() == std::vector<int>({0, 2, 3, 5}), "row_ptr");
  require(csr.col_idx() == std::vector<int>({0, 3, 1, 0, 3}), "col_idx");
  require(csr.values() == std::vector<double>({1.0, 3.0, 2.0, 5.0, 7.0}), "values");
}

void csr_to_csc_preserves_entries() {
  const CscMatrix csc = CsrMatrix(3, 4, {0, 2, 3, 5}, {0, 3, 1, 0, 3}, {1.0, 3.0, 2.0, 5.0, 7.0}).to_csc();
  require(csc.rows() == 3 && csc.cols() == 4, "shape");
  require(csc.col_ptr() == std::vector<int>({0, 2, 3, 3, 5}), "col_ptr");
  require(csc.row_idx() == std::vector<int>({0, 2, 1, 0, 2}), "row_idx");
  require(csc.values() == std::vector<double>({1.0, 5.0, 2.0, 3.0, 7.0}), "values");
}

void lp_model_rejects_dimension_mismatch() {
  const CscMatrix csc(2, 2, {0, 1, 2}, {0, 1}, {1.0, 2.0});
  require(throws([&] { LpModel m(csc, {1.0}, {0.0, 0.0}, {1.0, 1.0}, {0.0, 0.0}, {10.0, 10.0}, 0.0); }),
          "mismatched objective accepted");
}

void lp_model_builds_both_views() {
  const LpModel m(CscMatrix(2, 3, {0, 1, 2, 3}, {0, 1, 0}, {1.0, 2.0, 3.0}), {4.0, 5.0, 6.0}, {0.0, -1.0},
                  {10.0, 12.0}, {0.0, 0.0, 0.0}, {100.0, 100.0, 100.0}, 7.0);
  require(m.num_rows() == 2 && m.num_cols() == 3, "shape");
  require(m.csr().row_ptr() == std::vector<int>({0, 2, 3}), "csr row_ptr");
  require(std::fabs(m.obj_constant() - 7.0) <= 1e-12, "obj_constant");
}

void parses_rows_columns_rhs_and_bounds() {
  const auto model = parse(
      "NAME          toy\nROWS\n N  COST\n E  EQ1\n G  GE1\n L  LE1\nCOLUMNS\n"
      "    X1        COST       3.0   EQ1        2.0\n    X1        GE1        1.0\n"
      "    X2        COST       5.0   LE1       -4.0\n    X2        EQ1        1.0\nRHS\n"
      "    RHS1      EQ1       10.0   GE1        7.0\n    RHS1      LE1       20.0\nBOUNDS\n"
      " LO BND1      X1        -1.0\n UP BND1      X1         6.0\n FR BND1      X2\nENDATA\n");
  const auto& lp = model.lp;
  require(model.name == "toy" && lp.num_rows() == 3 && lp.num_cols() == 2, "shape");
  require(lp.objective() == std::vector<double>({3.0, 5.0}), "objective");
  require(lp.row_lower() == std::vector<double>({10.0, 7.0, -kInf}), "row_lower");
  require(lp.row_upper() == std::vector<double>({10.0, kInf, 20.0}), "row_upper");
  require(lp.col_lower() == std::vector<double>({-1.0, -kInf}), "col_lower");
  require(lp.col_upper() == std::vector<double>({6.0, kInf}), "col_upper");
  require(lp.csc().col_ptr() == std::vector<int>({0, 2, 4}), "col_ptr");
  require(lp.csc().row_idx() == std::vector<int>({0, 1, 0, 2}), "row_idx");
  require(lp.csc().values() == std::vector<double>({2.0, 1.0, 1.0, -4.0}), "values");
}

void supports_ranges_and_duplicate_entries() {
  const auto model = parse(
      "NAME ranged\nROWS\n N OBJ\n E EQ1\n L LE1\n G GE1\nCOLUMNS\n"
      "    X1        OBJ        1.0   EQ1        1.0\n    X1        LE1        2.0   LE1        3.0\n"
      "    X2        OBJ        2.0   GE1       -1.0\nRHS\n    RHS1      EQ1       10.0   LE1       20.0\n"
      "    RHS1      GE1        5.0\nRANGES\n    RNG1      EQ1        4.0   LE1        6.0\n"
      "    RNG1      GE1        7.0\nBOUNDS\n UP BND1      X1         1.0\n UP BND1      X2         9.0\nENDATA\n");
  const auto& lp = model.lp;
  require(model.name == "ranged" && lp.num_rows() == 3 && lp.num_cols() == 2, "shape");
  require(lp.objective() == std::vector<double>({1.0, 2.0}), "objective");
  require(lp.row_lower() == std::vector<double>({10.0, 14.0, 5.0}), "row_lower");
  require(lp.row_upper() == std::vector<double>({14.0, 20.0, 12.0}), "row_upper");
  require(lp.col_lower() == std::vector<double>({0.0, 0.0}), "col_lower");
  require(lp.col_upper() == std::vector<double>({1.0, 9.0}), "col_upper");
  require(lp.csc().col_ptr() == std::vector<int>({0, 2, 3}), "col_ptr");
  require(lp.csc().row_idx() == std::vector<int>({0, 1, 2}), "row_idx");
  require(lp.csc().values() == std::vector<double>({1.0, 5.0, -1.0}), "values");
}

void micro_bound_reads_above_its_decimal() {
  const auto model = parse(
      "NAME micro\nROWS\n N OBJ\n L ROW1\nCOLUMNS\n    X1        OBJ        1.0   ROW1       1.0\nRHS\n"
      "    RHS1      ROW1       1.0\nBOUNDS\n UP BND1      X1         0.000001\nENDATA\n");
  require(model.lp.col_upper().size() == 1 && model.lp.col_upper()[0] > 1.0e-6, "micro bound");
}

void falls_back_to_fixed_format() {
  const auto model = parse(
      "NAME          fixed\nROWS\n N  COST\n L  ROW00001\nCOLUMNS\n    COL00001  COST       1.0   ROW00001   2.0\n"
      "RHS\n    RHS00001  ROW00001   3.0\nENDATA\n");
  const auto& lp = model.lp;
  require(model.name == "fixed" && lp.num_rows() == 1 && lp.num_cols() == 1, "shape");
  require(lp.objective() == std::vector<double>({1.0}), "objective");
  require(lp.row_upper() == std::vector<double>({3.0}), "row_upper");
  require(lp.csc().values() == std::vector<double>({2.0}), "values");
}

void validates_objective_sense() {
  for (const std::string sense : {"OBJSENSE MIN\n", "OBJSENSE\n MIN\n"}) {
    std::string t = numeric_model();
    t.insert(t.find("ROWS"), sense);
    require(parse(t).lp.objective()[0] == 1.0, "MIN accepted");
  }
  for (const std::string sense : {"OBJSENSE MAX\n", "OBJSENSE\n MAX\n", "OBJSENSE INVALID\n", "OBJSENSE\n INVALID\n",
                                  "OBJSENSE\n", "OBJSENSE MIN MAX\n"}) {
    std::string t = numeric_model();
    t.insert(t.find("ROWS"), sense);
    require(rejected(t), "bad OBJSENSE accepted");
  }
}

void rejects_nonfinite_numeric_data() {
  for (const std::string v : {"nan", "NaN", "inf", "-inf", "1e999", "1e99999999999999"}) {
    require(rejected(numeric_model("    X OBJ " + v + " R 1\n")), "objective");
    require(rejected(numeric_model("    X OBJ 1 R " + v + "\n")), "matrix");
    require(rejected(numeric_model("    X OBJ 1 R 1\n", "    RHS1 OBJ " + v + " R 1\n")), "rhs objective");
    require(rejected(numeric_model("    X OBJ 1 R 1\n", "    RHS1 R " + v + "\n")), "rhs");
    std::string ranged = numeric_model();
    ranged.insert(ranged.find("BOUNDS"), "RANGES\n    RNG1 R " + v + "\n");
    require(rejected(ranged), "range");
  }
  for (const std::string b : {"LO BND1 X nan", "UP BND1 X nan", "FX BND1 X nan", "LO BND1 X inf", "UP BND1 X -inf",
                              "FX BND1 X inf", "UP BND1 X 1e999"})
    require(rejected(numeric_model("    X OBJ 1 R 1\n", "    RHS1 R 1\n", " " + b + "\n")), "bound");
  require(rejected(numeric_model("    X OBJ 1e308 OBJ 1e308\n    X R 1\n")), "objective sum overflow");
  require(rejected(numeric_model("    X OBJ 1 R 1e308\n    X R 1e308\n")), "matrix sum overflow");
  std::string ranged = numeric_model("    X OBJ 1 R 1\n", "    RHS1 R -1e308\n");
  ranged.insert(ranged.find("BOUNDS"), "RANGES\n    RNG1 R 1e308\n");
  require(rejected(ranged), "range overflow");
}

void preserves_unbounded_variables() {
  for (const std::string b : {" FR BND1 X\n", " FR BND1 X 0\n", " MI BND1 X\n PL BND1 X\n",
                              " LO BND1 X -inf\n UP BND1 X +Infinity\n"}) {
    const auto m = parse(numeric_model("    X OBJ 1 R 1\n", "    RHS1 R 1\n", b));
    require(m.lp.col_lower()[0] == -kInf && m.lp.col_upper()[0] == kInf, "free column");
  }
}

void rejects_incomplete_records() {
  require(rejected(numeric_model("    X OBJ 1 R\n")), "columns");
  require(rejected(numeric_model("    X OBJ 1 R 1\n", "    RHS1 R 1 OBJ\n")), "rhs");
  std::string t = numeric_model();
  t.insert(t.find("BOUNDS"), "RANGES\n    RNG1 R 1 R\n");
  require(rejected(t), "ranges");
  for (const std::string b : {" UP BND1 X\n", " UP BND1 X 2 EXTRA\n", " FR BND1 X 0 EXTRA\n"})
    require(rejected(numeric_model("    X OBJ 1 R 1\n", "    RHS1 R 1\n", b)), "bounds");
}

void rejects_integer_models() {
  for (const std::string b : {"BV BND1 X", "LI BND1 X 0", "UI BND1 X 1", "SI BND1 X 1"})
    require(rejected(numeric_model("    X OBJ 1 R 1\n", "    RHS1 R 0.5\n", " " + b + "\n")), "integer bound");
  for (const std::string mk : {"    MARK0000 'MARKER' 'INTORG'\n", "    MARK0000 MARKER INTORG\n"})
    require(rejected(numeric_model(mk + "    X OBJ 1 R 1\n")), "marker");
}

void preserves_extreme_finite_numbers() {
  const std::pair<const char*, double> cases[] = {{"1e-309", 1e-309},
                                                  {"-1D-309", -1e-309},
                                                  {"1e-320", 1e-320},
                                                  {"4.9406564584124654e-324", std::numeric_limits<double>::denorm_min()},
                                                  {"0.000000001e309", 1e300},
                                                  {"0.1e-308", 1e-309},
                                                  {"0e999", 0.0}};
  for (const auto& [tok, want] : cases) {
    const double got = parse(numeric_model(std::string("    X OBJ ") + tok + " R 1\n")).lp.objective()[0];
    require(std::isfinite(got) && got == want, tok);
  }
  for (const std::string tok : {"1e309", "1e-999", "1e99999999999", "1e-99999999999"})
    require(rejected(numeric_model("    X OBJ " + tok + " R 1\n")), "out of range number");
}

}  // namespace

int main() {
  const std::vector<std::pair<const char*, std::function<void()>>> cases = {
      {"csr_rejects_bad_row_pointer_size", csr_rejects_bad_row_pointer_size},
      {"csc_to_csr_preserves_entries", csc_to_csr_preserves_entries},
      {"csr_to_csc_preserves_entries", csr_to_csc_preserves_entries},
      {"lp_model_rejects_dimension_mismatch", lp_model_rejects_dimension_mismatch},
      {"lp_model_builds_both_views", lp_model_builds_both_views},
      {"parses_rows_columns_rhs_and_bounds", parses_rows_columns_rhs_and_bounds},
      {"supports_ranges_and_duplicate_entries", supports_ranges_and_duplicate_entries},
      {"micro_bound_reads_above_its_decimal", micro_bound_reads_above_its_decimal},
      {"falls_back_to_fixed_format", falls_back_to_fixed_format},
      {"validates_objective_sense", validates_objective_sense},
      {"rejects_nonfinite_numeric_data", rejects_nonfinite_numeric_data},
      {"preserves_unbounded_variables", preserves_unbounded_variables},
      {"rejects_incomplete_records", rejects_incomplete_records},
      {"rejects_integer_models", rejects_integer_models},
      {"preserves_extreme_finite_numbers", preserves_extreme_finite_numbers},
  };
  for (const auto& [name, fn] : cases) {
    try {
      fn();
    } catch (const std::exception& e) {
      std::printf("FAIL %s: %s\n", name, e.what());
      return 1;
    }
    std::printf("ok %s\n", name);
  }
  return 0;
}
