// Continuous-MPS reader: matrices, LP model and parser (interface in
// include/mps_reader/).  Its decisions follow the reference reader
// (cpp/mps_reader/src/io/mpsreader.cpp:1-760, src/core/*.cpp) so that the
// same file gives the same LP bit for bit:
//   * free format first; on any error the fixed column layout (fields at
//     columns 2-3, 5-12, 15-22, 25-36, 40-47, 50-61; NAME from column 15),
//     and the free-format error is reported if both fail (:727-755);
//   * numbers by the reader's own decimal accumulation (:200-301): digits
//     into value*10+d, fraction digits as d*s with s = 0.1, 0.01, ...
//     (each step s *= 0.1), a decimal exponent applied as one multiply or
//     divide by pow(10, e); strtod (with D/d read as E) when there are more
//     than 18 digits, the exponent exceeds 308 or the scaled value leaves the
//     normal range.  The accumulation is not correctly rounded (0.000001
//     reads as slightly more than 1e-6), and the LP's values depend on it;
//   * rows in ROWS order, columns in order of first appearance; repeated
//     (row, column) entries and repeated objective entries are summed in
//     file order; an RHS entry on the objective row sets the objective
//     constant to minus its value; only the first RHS, RANGES and BOUNDS
//     set names are read (:378-401, :423-437, :439-498);
//   * RANGES on E rows widen the upper side for r >= 0 and the lower side
//     otherwise; on L rows lower = upper - |r|; on G rows upper = lower + |r|
//     (:403-421);
//   * a column with no bound is [0, inf); one with only an upper bound has
//     lower -inf when that bound is negative, else 0 (:500-514).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <istream>
#include <limits>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "mps_reader/mpsreader.hpp"

namespace mps_reader {

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

void check_compressed(const char* what, int major, int minor, const std::vector<int>& ptr,
                      const std::vector<int>& idx, const std::vector<double>& val) {
  const std::string w(what);
  if (major < 0 || minor < 0) throw std::invalid_argument(w + ": negative dimension");
  if (ptr.size() != static_cast<size_t>(major) + 1) throw std::invalid_argument(w + ": pointer size != major + 1");
  if (idx.size() != val.size()) throw std::invalid_argument(w + ": index and value sizes differ");
  if (ptr.front() != 0) throw std::invalid_argument(w + ": pointers must start at zero");
  for (size_t k = 1; k < ptr.size(); ++k)
    if (ptr[k] < ptr[k - 1]) throw std::invalid_argument(w + ": pointers must not decrease");
  if (ptr.back() != static_cast<int>(val.size())) throw std::invalid_argument(w + ": last pointer != nnz");
  for (int i : idx)
    if (i < 0 || i >= minor) throw std::invalid_argument(w + ": index out of range");
}

// Counting-sort transposition of a compressed matrix (stable).
void transpose_compressed(int major, int minor, const std::vector<int>& ptr, const std::vector<int>& idx,
                          const std::vector<double>& val, std::vector<int>& tptr, std::vector<int>& tidx,
                          std::vector<double>& tval) {
  tptr.assign(static_cast<size_t>(minor) + 1, 0);
  for (int i : idx) ++tptr[static_cast<size_t>(i) + 1];
  for (int k = 0; k < minor; ++k) tptr[k + 1] += tptr[k];
  std::vector<int> cursor(tptr.begin(), tptr.end() - 1);
  tidx.assign(idx.size(), 0);
  tval.assign(val.size(), 0.0);
  for (int a = 0; a < major; ++a)
    for (int p = ptr[a]; p < ptr[a + 1]; ++p) {
      const int dst = cursor[idx[p]]++;
      tidx[dst] = a;
      tval[dst] = val[p];
    }
}

}  // namespace

CsrMatrix::CsrMatrix(int rows, int cols, std::vector<int> row_ptr, std::vector<int> col_idx,
                     std::vector<double> values)
    : rows_(rows), cols_(cols), row_ptr_(std::move(row_ptr)), col_idx_(std::move(col_idx)),
      values_(std::move(values)) {
  check_compressed("CSR", rows_, cols_, row_ptr_, col_idx_, values_);
}

CscMatrix CsrMatrix::to_csc() const {
  std::vector<int> p, i;
  std::vector<double> v;
  transpose_compressed(rows_, cols_, row_ptr_, col_idx_, values_, p, i, v);
  return CscMatrix(rows_, cols_, std::move(p), std::move(i), std::move(v));
}

CscMatrix::CscMatrix(int rows, int cols, std::vector<int> col_ptr, std::vector<int> row_idx,
                     std::vector<double> values)
    : rows_(rows), cols_(cols), col_ptr_(std::move(col_ptr)), row_idx_(std::move(row_idx)),
      values_(std::move(values)) {
  check_compressed("CSC", cols_, rows_, col_ptr_, row_idx_, values_);
}

CsrMatrix CscMatrix::to_csr() const {
  std::vector<int> p, i;
  std::vector<double> v;
  transpose_compressed(cols_, rows_, col_ptr_, row_idx_, values_, p, i, v);
  return CsrMatrix(rows_, cols_, std::move(p), std::move(i), std::move(v));
}

LpModel::LpModel(CscMatrix matrix, std::vector<double> objective, std::vector<double> row_lower,
                 std::vector<double> row_upper, std::vector<double> col_lower, std::vector<double> col_upper,
                 double obj_constant)
    : csc_(std::move(matrix)), csr_(csc_.to_csr()), objective_(std::move(objective)),
      row_lower_(std::move(row_lower)), row_upper_(std::move(row_upper)), col_lower_(std::move(col_lower)),
      col_upper_(std::move(col_upper)), obj_constant_(obj_constant) {
  const size_t m = static_cast<size_t>(num_rows()), n = static_cast<size_t>(num_cols());
  if (objective_.size() != n || col_lower_.size() != n || col_upper_.size() != n || row_lower_.size() != m ||
      row_upper_.size() != m)
    throw std::invalid_argument("LpModel: vector sizes do not match the matrix");
  for (size_t i = 0; i < m; ++i)
    if (row_lower_[i] > row_upper_[i])
      throw std::invalid_argument("LpModel: row_lower[" + std::to_string(i) + "] exceeds row_upper");
  for (size_t j = 0; j < n; ++j)
    if (col_lower_[j] > col_upper_[j])
      throw std::invalid_argument("LpModel: col_lower[" + std::to_string(j) + "] exceeds col_upper");
}

int LpModel::num_rows() const noexcept { return csc_.rows(); }
int LpModel::num_cols() const noexcept { return csc_.cols(); }
const CscMatrix& LpModel::csc() const { return csc_; }
const CsrMatrix& LpModel::csr() const { return csr_; }

namespace {

bool space(char ch) { return std::isspace(static_cast<unsigned char>(ch)) != 0; }

std::string trimmed(const std::string& s, size_t b, size_t e) {
  if (b >= s.size()) return {};
  e = std::min(e, s.size());
  while (b < e && space(s[b])) ++b;
  while (e > b && space(s[e - 1])) --e;
  return s.substr(b, e - b);
}

std::vector<std::string> words(const std::string& s) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < s.size()) {
    while (i < s.size() && space(s[i])) ++i;
    const size_t b = i;
    while (i < s.size() && !space(s[i])) ++i;
    if (i > b) out.push_back(s.substr(b, i - b));
  }
  return out;
}

// A data or header line; comments ('*' or '&' as the first non-blank
// character) and blank lines have no fields.
struct Card {
  int line = 0;
  bool header = false;
  std::vector<std::string> f;
};

Card make_card(const std::string& s, int line, bool fixed) {
  Card c;
  c.line = line;
  size_t first = 0;
  while (first < s.size() && space(s[first])) ++first;
  if (first == s.size() || s[first] == '*' || s[first] == '&') return c;
  c.header = !space(s[0]);
  if (c.header || !fixed) {
    c.f = words(s);
    if (c.header && fixed && !c.f.empty() && c.f[0] == "NAME") {
      const std::string nm = trimmed(s, 14, s.size());
      if (!nm.empty()) c.f = {"NAME", nm};
    } else if (c.header && c.f.size() >= 2 && c.f[0] == "OBJECT" && c.f[1] == "BOUND") {
      c.f = {"OBJECT BOUND"};
    }
    return c;
  }
  static const size_t cols[6][2] = {{1, 3}, {4, 12}, {14, 22}, {24, 36}, {39, 47}, {49, 61}};
  for (const auto& fc : cols) {
    std::string t = trimmed(s, fc[0], fc[1]);
    if (!t.empty()) c.f.push_back(std::move(t));
  }
  return c;
}

double via_strtod(const std::string& tok) {
  std::string t = tok;
  for (char& ch : t)
    if (ch == 'd' || ch == 'D') ch = 'E';
  char* end = nullptr;
  errno = 0;
  const double v = std::strtod(t.c_str(), &end);
  if (end == t.c_str() || *end != '\0') throw std::invalid_argument("invalid numeric token in MPS: " + tok);
  if (errno == ERANGE && (v == 0.0 || !std::isfinite(v)))
    throw std::out_of_range("numeric token out of double range in MPS: " + tok);
  return v;
}

bool digit(char ch) { return ch >= '0' && ch <= '9'; }

double number(const std::string& tok) {
  const size_t n = tok.size();
  size_t i = 0;
  bool neg = false;
  if (i < n && (tok[i] == '-' || tok[i] == '+')) neg = tok[i++] == '-';
  double v = 0.0, scale = 0.1;
  int nd = 0;
  for (; i < n && digit(tok[i]); ++i) {
    v = v * 10.0 + static_cast<double>(tok[i] - '0');
    if (++nd > 18) return via_strtod(tok);
  }
  if (i < n && tok[i] == '.') {
    for (++i; i < n && digit(tok[i]); ++i) {
      v += static_cast<double>(tok[i] - '0') * scale;
      scale *= 0.1;
      if (++nd > 18) return via_strtod(tok);
    }
  }
  if (nd == 0) return via_strtod(tok);
  if (i < n) {
    if (tok[i] != 'e' && tok[i] != 'E' && tok[i] != 'd' && tok[i] != 'D') return via_strtod(tok);
    ++i;
    bool eneg = false;
    if (i < n && (tok[i] == '-' || tok[i] == '+')) eneg = tok[i++] == '-';
    int e = 0, ed = 0;
    for (; i < n; ++i, ++ed) {
      if (!digit(tok[i])) return via_strtod(tok);
      const int d = tok[i] - '0';
      if (e > (std::numeric_limits<int>::max() - d) / 10) return via_strtod(tok);
      e = 10 * e + d;
    }
    if (ed == 0 || e > 308) return via_strtod(tok);
    const double p = std::pow(10.0, e);
    const double scaled = eneg ? v / p : v * p;
    if (!std::isfinite(scaled) || (v != 0.0 && std::fabs(scaled) < std::numeric_limits<double>::min()))
      return via_strtod(tok);
    v = scaled;
  }
  return neg ? -v : v;
}

double finite(double v, const std::string& what) {
  if (!std::isfinite(v)) throw std::invalid_argument("non-finite value in MPS: " + what);
  return v;
}

// A value field; in a bound, "inf"/"infinity" (any case, optional sign) read
// as infinite, but a numeric overflow does not.
double value(const std::string& tok, bool bound = false) {
  const double v = number(tok);
  if (bound && std::isinf(v)) {
    std::string t;
    for (char ch : tok) t.push_back(static_cast<char>(std::tolower(static_cast<unsigned char>(ch))));
    if (!t.empty() && (t[0] == '+' || t[0] == '-')) t.erase(0, 1);
    if (t == "inf" || t == "infinity") return v;
  }
  return finite(v, tok);
}

enum class Sec { kNone, kObjSense, kRows, kColumns, kRhs, kBounds, kRanges, kQuadratic };

struct Row {
  char type;
  int index;  // -1: a free (N) row
};

class Parser {
 public:
  std::string name;
  std::string obj_row;
  std::unordered_map<std::string, Row> rows;
  std::unordered_map<std::string, int> cols;
  std::vector<char> row_type;
  std::vector<std::vector<std::pair<int, double>>> col_entries;
  std::vector<double> c, rl, ru, cl, cu;
  double obj_constant = 0.0;

  void parse(const std::vector<std::string>& lines, bool fixed) {
    Sec sec = Sec::kNone;
    bool want_sense = false;
    for (size_t k = 0; k < lines.size(); ++k) {
      const Card card = make_card(lines[k], static_cast<int>(k + 1), fixed && sec != Sec::kObjSense);
      if (card.f.empty()) continue;
      if (card.header) {
        if (want_sense) throw std::invalid_argument("missing MPS OBJSENSE value");
        const std::string& h = card.f[0];
        if (h == "NAME") {
          if (card.f.size() >= 2) name = card.f[1];
          continue;
        }
        if (h == "ENDATA") break;
        sec = section(h);
        if (sec == Sec::kQuadratic) throw std::invalid_argument("quadratic MPS sections are not supported");
        if (sec == Sec::kObjSense) {
          if (card.f.size() == 2) sense(card.f[1]);
          else if (card.f.size() == 1) want_sense = true;
          else throw std::invalid_argument("invalid MPS OBJSENSE entry");
        }
        continue;
      }
      switch (sec) {
        case Sec::kObjSense:
          if (!want_sense || card.f.size() != 1) throw std::invalid_argument("invalid MPS OBJSENSE entry");
          sense(card.f[0]);
          want_sense = false;
          break;
        case Sec::kRows: on_row(card); break;
        case Sec::kColumns:
          if (card.f.size() >= 2 && (card.f[1] == "'MARKER'" || card.f[1] == "MARKER"))
            throw std::invalid_argument("MPS integer markers are not supported, line " + std::to_string(card.line));
          on_column(card);
          break;
        case Sec::kRhs: on_rhs(card); break;
        case Sec::kRanges: on_range(card); break;
        case Sec::kBounds: on_bound(card); break;
        default: throw std::invalid_argument("MPS data entry outside a supported section");
      }
    }
    if (want_sense) throw std::invalid_argument("missing MPS OBJSENSE value");
    if (row_type.empty() && cols.empty()) throw std::invalid_argument("empty MPS model");
    for (size_t j = 0; j < cl.size(); ++j) {
      const bool lo = !std::isnan(cl[j]), up = !std::isnan(cu[j]);
      if (!lo) cl[j] = up && cu[j] < 0.0 ? -kInf : 0.0;
      if (!up) cu[j] = kInf;
    }
  }

 private:
  std::string rhs_set, range_set, bound_set;

  static Sec section(const std::string& h) {
    if (h == "OBJSENSE") return Sec::kObjSense;
    if (h == "ROWS") return Sec::kRows;
    if (h == "COLUMNS") return Sec::kColumns;
    if (h == "RHS") return Sec::kRhs;
    if (h == "BOUNDS") return Sec::kBounds;
    if (h == "RANGES") return Sec::kRanges;
    if (h == "QUADOBJ" || h == "QMATRIX") return Sec::kQuadratic;
    throw std::invalid_argument("unknown MPS section header: " + h);
  }

  static void sense(const std::string& s) {
    if (s == "MAX") throw std::invalid_argument("MPS OBJSENSE MAX is not supported");
    if (s != "MIN") throw std::invalid_argument("unsupported MPS OBJSENSE: " + s);
  }

  // Pairs (name, value) from field `from` on: an odd total field count.
  static void need_pairs(const Card& card, const char* sec) {
    if (card.f.size() < 3 || card.f.size() % 2 != 1)
      throw std::invalid_argument(std::string("invalid ") + sec + " entry at line " + std::to_string(card.line));
  }

  const Row& row_of(const std::string& nm, const char* sec) const {
    const auto it = rows.find(nm);
    if (it == rows.end()) throw std::invalid_argument(std::string("unknown MPS row in ") + sec + ": " + nm);
    return it->second;
  }

  void on_row(const Card& card) {
    if (card.f.size() < 2) throw std::invalid_argument("invalid ROWS entry at line " + std::to_string(card.line));
    const char t = card.f[0].empty() ? '\0' : card.f[0][0];
    const std::string& nm = card.f[1];
    if (t == 'N') {
      if (obj_row.empty()) obj_row = nm;
      rows.emplace(nm, Row{'N', -1});
      return;
    }
    if (t != 'E' && t != 'G' && t != 'L') throw std::invalid_argument("unsupported MPS row type: " + std::string(1, t));
    rows.emplace(nm, Row{t, static_cast<int>(row_type.size())});
    row_type.push_back(t);
    rl.push_back(t == 'L' ? -kInf : 0.0);
    ru.push_back(t == 'G' ? kInf : 0.0);
  }

  int column(const std::string& nm) {
    const auto it = cols.find(nm);
    if (it != cols.end()) return it->second;
    const int j = static_cast<int>(cols.size());
    cols.emplace(nm, j);
    col_entries.emplace_back();
    c.push_back(0.0);
    cl.push_back(std::numeric_limits<double>::quiet_NaN());
    cu.push_back(std::numeric_limits<double>::quiet_NaN());
    return j;
  }

  void on_column(const Card& card) {
    need_pairs(card, "COLUMNS");
    const int j = column(card.f[0]);
    for (size_t k = 1; k + 1 < card.f.size(); k += 2) {
      const double v = value(card.f[k + 1]);
      const Row& r = row_of(card.f[k], "COLUMNS");
      if (r.index >= 0) {
        col_entries[static_cast<size_t>(j)].emplace_back(r.index, v);
      } else if (card.f[k] == obj_row) {
        c[static_cast<size_t>(j)] = finite(c[static_cast<size_t>(j)] + v, "summed objective coefficient");
      }
    }
  }

  void on_rhs(const Card& card) {
    need_pairs(card, "RHS");
    if (rhs_set.empty()) rhs_set = card.f[0];
    if (card.f[0] != rhs_set) return;
    for (size_t k = 1; k + 1 < card.f.size(); k += 2) {
      const Row& r = row_of(card.f[k], "RHS");
      if (r.index < 0) {
        if (card.f[k] == obj_row) obj_constant = -value(card.f[k + 1]);
        continue;
      }
      const double v = value(card.f[k + 1]);
      const size_t i = static_cast<size_t>(r.index);
      if (r.type != 'L') rl[i] = v;
      if (r.type != 'G') ru[i] = v;
    }
  }

  void on_range(const Card& card) {
    need_pairs(card, "RANGES");
    if (range_set.empty()) range_set = card.f[0];
    if (card.f[0] != range_set) return;
    for (size_t k = 1; k + 1 < card.f.size(); k += 2) {
      const double v = value(card.f[k + 1]);
      const Row& r = row_of(card.f[k], "RANGES");
      if (r.index < 0) continue;
      const size_t i = static_cast<size_t>(r.index);
      if (r.type == 'E') {
        if (v >= 0.0) ru[i] = finite(ru[i] + v, "ranged row bound");
        else rl[i] = finite(rl[i] + v, "ranged row bound");
      } else if (r.type == 'L') {
        rl[i] = finite(ru[i] - std::fabs(v), "ranged row bound");
      } else {
        ru[i] = finite(rl[i] + std::fabs(v), "ranged row bound");
      }
    }
  }

  void on_bound(const Card& card) {
    if (card.f.size() < 3) throw std::invalid_argument("invalid BOUNDS entry at line " + std::to_string(card.line));
    const std::string& t = card.f[0];
    if (t == "BV" || t == "LI" || t == "UI" || t == "SI")
      throw std::invalid_argument("integer MPS bounds are not supported, line " + std::to_string(card.line));
    const bool valueless = t == "FR" || t == "MI" || t == "PL";
    if (card.f.size() > 4 || (!valueless && card.f.size() != 4))
      throw std::invalid_argument("invalid BOUNDS field count at line " + std::to_string(card.line));
    if (bound_set.empty()) bound_set = card.f[1];
    if (card.f[1] != bound_set) return;
    const auto it = cols.find(card.f[2]);
    if (it == cols.end()) throw std::invalid_argument("unknown MPS column in BOUNDS: " + card.f[2]);
    const size_t j = static_cast<size_t>(it->second);
    if (valueless) {
      if (t != "PL") cl[j] = -kInf;
      if (t != "MI") cu[j] = kInf;
      return;
    }
    const double v = value(card.f[3], true);
    if ((t == "LO" && v == kInf) || (t == "UP" && v == -kInf) || (t == "FX" && !std::isfinite(v)))
      throw std::invalid_argument("invalid infinite MPS bound: " + t);
    if (t == "LO") {
      cl[j] = v;
    } else if (t == "UP") {
      cu[j] = v;
    } else if (t == "FX") {
      cl[j] = v;
      cu[j] = v;
    } else {
      throw std::invalid_argument("unsupported MPS bound type: " + t);
    }
  }
};

MpsModel assemble(Parser&& p) {
  std::vector<int> ptr{0}, idx;
  std::vector<double> val;
  for (auto& e : p.col_entries) {
    std::stable_sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (size_t b = 0; b < e.size();) {
      size_t q = b;
      double s = 0.0;
      for (; q < e.size() && e[q].first == e[b].first; ++q) s += e[q].second;
      idx.push_back(e[b].first);
      val.push_back(finite(s, "summed matrix coefficient"));
      b = q;
    }
    ptr.push_back(static_cast<int>(val.size()));
  }
  const int m = static_cast<int>(p.row_type.size()), n = static_cast<int>(p.col_entries.size());
  LpModel lp(CscMatrix(m, n, std::move(ptr), std::move(idx), std::move(val)), std::move(p.c), std::move(p.rl),
             std::move(p.ru), std::move(p.cl), std::move(p.cu), p.obj_constant);
  return MpsModel{std::move(p.name), std::move(lp)};
}

}  // namespace

MpsModel read_mps(std::istream& input) {
  std::vector<std::string> lines;
  for (std::string s; std::getline(input, s);) lines.push_back(std::move(s));
  try {
    Parser p;
    p.parse(lines, false);
    return assemble(std::move(p));
  } catch (const std::exception& free_error) {
    const std::string msg = free_error.what();
    try {
      Parser p;
      p.parse(lines, true);
      return assemble(std::move(p));
    } catch (const std::exception&) {
      throw std::invalid_argument(msg);
    }
  }
}

}  // namespace mps_reader
