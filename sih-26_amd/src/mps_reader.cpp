// SPDX-License-Identifier: Apache-2.0
// MPS model reader (reference interface: include/sankhya/io.hpp:19-53, sankhya::io::read_mps,
// parse_mps_format and read_model; behaviour of src/io/mps_reader.cpp and src/io/mps_bounds.cpp,
// which follow IBM's MPS description).
//
// One pass over the file collects rows, columns, right-hand sides, ranges and bounds; row bounds
// are resolved afterwards from each row's (type, rhs, range), and the matrix is built through
// SparseMatrix (CSC, rows ascending per column, explicit zeros and entries under 1e-11 dropped). Free MPS splits lines on blanks and ends a line at a token
// opening with '$'; fixed MPS reads the six fields at their columns. Section headers start in
// column 1 in both dialects. With MpsFormat::kAuto a file is read free, and fixed only when the
// free reading fails (names with embedded blanks give lines of impossible shape).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#include <zlib.h>

#include "sankhya/io.hpp"
#include "text_format.hpp"

namespace sankhya::io {
namespace {

using hip_engine::text::format;

enum class Section { kNone, kObjsenseValue, kRows, kColumns, kRhs, kRanges, kBounds };

constexpr Index kObjective = -1;  // row handle of the objective row
constexpr Index kDropped = -2;    // row handle of an N row after the objective
constexpr Index kUnknown = -3;

std::string upper(std::string s) {
  for (char& c : s) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
  return s;
}

bool blank(char c) { return std::isspace(static_cast<unsigned char>(c)) != 0; }

std::string strip_quotes(const std::string& s) {
  if (s.size() >= 2 && ((s.front() == '\'' && s.back() == '\'') || (s.front() == '"' && s.back() == '"')))
    return s.substr(1, s.size() - 2);
  return s;
}

std::vector<std::string> words(const std::string& line) {
  std::vector<std::string> out;
  std::size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && blank(line[i])) ++i;
    const std::size_t b = i;
    while (i < line.size() && !blank(line[i])) ++i;
    if (i > b) out.emplace_back(line.substr(b, i - b));
  }
  return out;
}

// Fixed MPS data fields: 1-based columns 2-3, 5-12, 15-22, 25-36, 40-47 and 50-61. Empty fields
// are skipped, so an absent vector name shifts the rest left exactly as in free MPS.
std::vector<std::string> fixed_fields(const std::string& line) {
  static constexpr std::size_t kAt[6] = {1, 4, 14, 24, 39, 49};
  static constexpr std::size_t kWidth[6] = {2, 8, 8, 12, 8, 12};
  std::vector<std::string> out;
  for (int f = 0; f < 6 && kAt[f] < line.size(); ++f) {
    const std::string field = line.substr(kAt[f], std::min(kWidth[f], line.size() - kAt[f]));
    const std::size_t b = field.find_first_not_of(" \t\r");
    if (b == std::string::npos) continue;
    out.push_back(field.substr(b, field.find_last_not_of(" \t\r") - b + 1));
  }
  return out;
}

// A number field: decimal or Fortran 'D' exponent; NaN and trailing characters are errors. An
// overflowing magnitude reads as +-inf; callers decide whether a field may be infinite.
bool parse_number(std::string text, double* out) {
  if (text.empty()) return false;
  for (char& c : text)
    if (c == 'D' || c == 'd') c = 'E';
  char* end = nullptr;
  const double v = std::strtod(text.c_str(), &end);
  if (end != text.c_str() + text.size() || std::isnan(v)) return false;
  *out = v;
  return true;
}

// The lines of a plain or gzip-compressed file (the reference reads `.gz` through zlib,
// src/io/line_reader.cpp:43-60); zlib passes an uncompressed file through as it is.
class LineSource {
 public:
  explicit LineSource(const std::string& path) : gz_(gzopen(path.c_str(), "rb")) {
    if (gz_ != nullptr) gzbuffer(gz_, 1 << 17);
  }
  ~LineSource() {
    if (gz_ != nullptr) gzclose(gz_);
  }
  LineSource(const LineSource&) = delete;
  LineSource& operator=(const LineSource&) = delete;
  bool open() const { return gz_ != nullptr; }
  // The next line without its newline; false at the end of the data or on a read error.
  bool next(std::string* line) {
    line->clear();
    char buf[4096];
    while (gzgets(gz_, buf, sizeof buf) != nullptr) {
      const std::size_t len = std::strlen(buf);
      line->append(buf, len);
      if (len > 0 && buf[len - 1] == '\n') {
        line->pop_back();
        return true;
      }
    }
    return !line->empty();
  }
  // Corrupt or truncated compressed data (a clean end of the data reads as no error).
  bool failed() const {
    int code = Z_OK;
    gzerror(gz_, &code);
    return code != Z_OK && code != Z_STREAM_END;
  }

 private:
  gzFile gz_;
};

class MpsReader {
 public:
  MpsReader(std::string path, MpsFormat dialect)
      : path_(std::move(path)), fixed_(dialect == MpsFormat::kFixed) {}

  ReadResult read(Model* model) {
    LineSource in(path_);
    if (!in.open()) return ReadResult::failure(format("{}: cannot open the file", path_));
    std::string line, err;
    bool ended = false;
    while (in.next(&line)) {
      ++line_no_;
      if (!line.empty() && line.back() == '\r') line.pop_back();
      if (line.empty() || line[0] == '*') continue;
      const bool header = !blank(line[0]);
      std::vector<std::string> tok = header || !fixed_ ? words(line) : fixed_fields(line);
      if (!fixed_) {
        for (std::size_t k = 0; k < tok.size(); ++k) {
          if (tok[k][0] == '$') {
            tok.resize(k);
            break;
          }
        }
      }
      if (tok.empty()) continue;
      if (header) {
        bool end = false;
        if (!section_header(tok, &end, &err)) return fail(err);
        if (end) {
          ended = true;
          break;
        }
        if (known_header_) continue;
        // An unindented line that names no section is data: free MPS files are not reliably
        // indented.
      }
      if (!data(tok, &err)) return fail(err);
    }
    if (!ended && in.failed())
      return ReadResult::failure(format("{}: corrupt or truncated compressed data after line {}",
                                        path_, line_no_));
    if (!ended) return ReadResult::failure(format("{}: the file ends without an ENDATA record", path_));
    return finish(model);
  }

 private:
  struct Entry {
    Index col, row;
    double value;
  };

  ReadResult fail(const std::string& what) const {
    return ReadResult::failure(format("{}:{}: {}", path_, line_no_, what));
  }

  void set_sense(const std::string& word) {
    const std::string w = upper(word);
    maximize_ = w == "MAX" || w == "MAXIMIZE";
  }

  bool section_header(const std::vector<std::string>& tok, bool* end, std::string* err) {
    const std::string key = upper(tok[0]);
    known_header_ = true;
    if (key == "NAME") {
      name_ = tok.size() > 1 ? tok[1] : std::string();
      section_ = Section::kNone;
    } else if (key == "OBJSENSE" || key == "OBJSENS") {
      if (tok.size() > 1) {
        set_sense(tok[1]);
        section_ = Section::kNone;
      } else {
        section_ = Section::kObjsenseValue;
      }
    } else if (key == "ROWS") {
      section_ = Section::kRows;
    } else if (key == "COLUMNS") {
      section_ = Section::kColumns;
    } else if (key == "RHS") {
      section_ = Section::kRhs;
    } else if (key == "RANGES") {
      section_ = Section::kRanges;
    } else if (key == "BOUNDS") {
      section_ = Section::kBounds;
    } else if (key == "QUADOBJ" || key == "QMATRIX" || key == "QSECTION" || key == "QUADS") {
      *err = format("{} holds a quadratic objective; this build reads linear programs only", tok[0]);
      return false;
    } else if (key == "ENDATA") {
      *end = true;
    } else {
      known_header_ = false;
    }
    return true;
  }

  bool data(const std::vector<std::string>& tok, std::string* err) {
    switch (section_) {
      case Section::kObjsenseValue:
        set_sense(tok[0]);
        section_ = Section::kNone;
        return true;
      case Section::kRows:
        return rows(tok, err);
      case Section::kColumns:
        return columns(tok, err);
      case Section::kRhs:
      case Section::kRanges:
        return rhs_or_ranges(tok, err);
      case Section::kBounds:
        return bounds(tok, err);
      case Section::kNone:
        break;
    }
    *err = "data line appears before any recognised section header";
    return false;
  }

  Index row_handle(const std::string& name) const {
    const auto it = row_index_.find(name);
    return it == row_index_.end() ? kUnknown : it->second;
  }

  bool rows(const std::vector<std::string>& tok, std::string* err) {
    // Exactly two fields: a fixed-format name with a blank gives three, which must fail so that
    // the automatic dialect choice retries in fixed columns.
    if (tok.size() != 2) {
      *err = format("ROWS entry has {} fields, expected exactly 2 (type and name)", tok.size());
      return false;
    }
    const std::string type = upper(tok[0]);
    if (type != "N" && type != "E" && type != "L" && type != "G") {
      *err = format("unknown row type '{}', expected N, L, G or E", tok[0]);
      return false;
    }
    if (row_index_.count(tok[1])) {
      *err = format("duplicate row name '{}'", tok[1]);
      return false;
    }
    Index handle;
    if (type == "N") {
      handle = have_objective_ ? kDropped : kObjective;  // later N rows constrain nothing
      have_objective_ = true;
    } else {
      handle = static_cast<Index>(row_type_.size());
      row_type_.push_back(type[0]);
      row_names_.push_back(tok[1]);
      rhs_.push_back(0.0);
      range_.push_back(0.0);
      has_range_.push_back(0);
    }
    row_index_.emplace(tok[1], handle);
    return true;
  }

  bool columns(const std::vector<std::string>& tok, std::string* err) {
    // 'MARKER' records switch integrality on (INTORG) and off (INTEND); writers put the quoted
    // keywords in different fields, so every token is looked at.
    bool marker = false, intorg = false, intend = false;
    for (const std::string& t : tok) {
      const std::string u = upper(strip_quotes(t));
      marker |= u == "MARKER";
      intorg |= u == "INTORG";
      intend |= u == "INTEND";
    }
    if (marker || intorg || intend) {
      if (!intorg && !intend) {
        *err = "MARKER record names neither INTORG nor INTEND";
        return false;
      }
      integer_ = intorg;
      return true;
    }
    if (tok.size() != 3 && tok.size() != 5) {
      *err = format("COLUMNS entry has {} fields, expected 3 or 5 (column, row, value[, row, value])",
                    tok.size());
      return false;
    }
    Index j;
    const auto it = col_index_.find(tok[0]);
    if (it != col_index_.end()) {
      j = it->second;
    } else {
      j = static_cast<Index>(col_names_.size());
      col_index_.emplace(tok[0], j);
      col_names_.push_back(tok[0]);
      cost_.push_back(0.0);
      lower_.push_back(0.0);  // [0, +inf) unless BOUNDS says otherwise, integer columns too
      upper_.push_back(kInfinity);
      lower_explicit_.push_back(0);
      integer_col_.push_back(integer_ ? 1 : 0);
    }
    for (std::size_t k = 1; k + 1 < tok.size(); k += 2) {
      const Index row = row_handle(tok[k]);
      if (row == kUnknown) {
        *err = format("COLUMNS entry names unknown row '{}'", tok[k]);
        return false;
      }
      double v = 0.0;
      if (!parse_number(tok[k + 1], &v)) {
        *err = format("'{}' is not a number", tok[k + 1]);
        return false;
      }
      if (row == kDropped) continue;
      // 1e30 means infinity in a bound, not in a coefficient; an overflowed one is a broken file.
      if (!std::isfinite(v)) {
        *err = format("coefficient for column '{}' is {}; a constraint or objective coefficient "
                      "must be finite", tok[0], v);
        return false;
      }
      if (row == kObjective) {
        if (!objective_seen_.insert(j).second) {
          *err = format("duplicate objective coefficient for column '{}'", tok[0]);
          return false;
        }
        cost_[static_cast<std::size_t>(j)] = v;
        continue;
      }
      const std::uint64_t key = (static_cast<std::uint64_t>(static_cast<std::uint32_t>(row)) << 32) |
                                static_cast<std::uint32_t>(j);
      if (!entry_seen_.insert(key).second) {
        *err = format("duplicate entry for row '{}' and column '{}'; MPS has no accumulate semantics",
                      tok[k], tok[0]);
        return false;
      }
      entries_.push_back({j, row, v});
    }
    return true;
  }

  // RHS and RANGES lines: [vector] (row, value) pairs. Only the first named vector is read.
  bool rhs_or_ranges(const std::vector<std::string>& tok, std::string* err) {
    const bool ranges = section_ == Section::kRanges;
    const char* what = ranges ? "RANGES" : "RHS";
    std::size_t k = 0;
    if (tok.size() % 2 == 1) {
      std::string& first = ranges ? range_vector_ : rhs_vector_;
      if (first.empty()) first = tok[0];
      if (tok[0] != first) return true;
      k = 1;
    }
    if (tok.size() <= k) {
      *err = format("{} entry has no row/value pair", what);
      return false;
    }
    for (; k + 1 < tok.size(); k += 2) {
      const Index row = row_handle(tok[k]);
      if (row == kUnknown) {
        *err = format("{} entry names unknown row '{}'", what, tok[k]);
        return false;
      }
      double v = 0.0;
      if (!parse_number(tok[k + 1], &v)) {
        *err = format("'{}' is not a number", tok[k + 1]);
        return false;
      }
      if (row == kDropped) continue;
      if (row == kObjective) {
        if (ranges) {
          *err = "RANGES entry on the objective row is not meaningful";
          return false;
        }
        objective_rhs_ = v;  // the negated objective constant
        continue;
      }
      const auto i = static_cast<std::size_t>(row);
      if (ranges) {
        range_[i] = v;
        has_range_[i] = 1;
      } else {
        rhs_[i] = v;
      }
    }
    return true;
  }

  bool bounds(const std::vector<std::string>& tok, std::string* err) {
    if (tok.size() < 2) {
      *err = "BOUNDS entry needs a type and a column";
      return false;
    }
    const std::string type = upper(tok[0]);
    const bool valued = type == "UP" || type == "LO" || type == "FX" || type == "LI" ||
                        type == "UI" || type == "SC";
    const bool valueless = type == "FR" || type == "MI" || type == "PL" || type == "BV";
    if (!valued && !valueless) {
      *err = format("unknown bound type '{}'; expected UP LO FX FR MI PL BV LI UI", tok[0]);
      return false;
    }
    // type [vector] column [value]: four fields name the vector for a valued type, three for a
    // valueless one (a stray value after a valueless type is ignored).
    std::size_t at;
    if (valued) {
      if (tok.size() != 3 && tok.size() != 4) {
        *err = format("bound type {} needs a column and a value", type);
        return false;
      }
      at = tok.size() == 4 ? 2 : 1;
    } else {
      at = tok.size() >= 3 ? 2 : 1;
    }
    if (at == 2) {
      if (bound_vector_.empty()) bound_vector_ = tok[1];
      if (tok[1] != bound_vector_) return true;
    }
    const auto it = col_index_.find(tok[at]);
    if (it == col_index_.end()) {
      *err = format("BOUNDS entry names column '{}', which has no COLUMNS entry", tok[at]);
      return false;
    }
    const auto j = static_cast<std::size_t>(it->second);
    double v = 0.0;
    if (valued) {
      if (!parse_number(tok[at + 1], &v)) {
        *err = format("'{}' is not a number", tok[at + 1]);
        return false;
      }
      v = normalize_infinity(v);
    }
    if (type == "SC") {
      *err = "semi-continuous bounds (SC) are not supported; the model would be misread as a "
             "plain integer problem";
      return false;
    }
    if (type == "UP" || type == "UI") {
      upper_[j] = v;
      if (type == "UI") integer_col_[j] = 1;
      // A negative upper bound on a column with no explicit lower bound makes the column
      // unbounded below; [0, v < 0] would be empty.
      if (v < 0.0 && !lower_explicit_[j]) lower_[j] = -kInfinity;
    } else if (type == "LO" || type == "LI") {
      lower_[j] = v;
      lower_explicit_[j] = 1;
      if (type == "LI") integer_col_[j] = 1;
    } else if (type == "FX") {
      lower_[j] = upper_[j] = v;
      lower_explicit_[j] = 1;
    } else if (type == "FR") {
      lower_[j] = -kInfinity;
      upper_[j] = kInfinity;
      lower_explicit_[j] = 1;
    } else if (type == "MI") {
      lower_[j] = -kInfinity;  // the upper bound is left alone
      lower_explicit_[j] = 1;
    } else if (type == "PL") {
      upper_[j] = kInfinity;
    } else {  // BV
      lower_[j] = 0.0;
      upper_[j] = 1.0;
      lower_explicit_[j] = 1;
      integer_col_[j] = 1;
    }
    return true;
  }

  ReadResult finish(Model* out) {
    Model md;
    md.name = name_;
    md.source_path = path_;
    md.sense = maximize_ ? ObjSense::kMaximize : ObjSense::kMinimize;
    md.objective_offset = normalize_zero(-objective_rhs_);
    const std::size_t m = row_type_.size();
    md.row_lower.resize(m);
    md.row_upper.resize(m);
    for (std::size_t i = 0; i < m; ++i) {
      // Without a range: L (-inf, b], G [b, +inf), E [b, b]. With range R: L [b - |R|, b],
      // G [b, b + |R|], E [b, b + R] for R >= 0 and [b + R, b] for R < 0.
      const double b = normalize_infinity(rhs_[i]);
      const double r = range_[i];
      double lo = b, hi = b;
      if (row_type_[i] == 'L') lo = has_range_[i] ? b - std::fabs(r) : -kInfinity;
      if (row_type_[i] == 'G') hi = has_range_[i] ? b + std::fabs(r) : kInfinity;
      if (row_type_[i] == 'E' && has_range_[i]) (r >= 0.0 ? hi : lo) = b + r;
      if (lo > hi)
        return ReadResult::failure(format("{}: row '{}' has an empty range [{}, {}]", path_,
                                          row_names_[i], lo, hi));
      md.row_lower[i] = lo;
      md.row_upper[i] = hi;
    }
    const std::size_t n = col_names_.size();
    md.col_names = col_names_;
    md.row_names = row_names_;
    md.col_cost = cost_;
    md.col_lower = lower_;
    md.col_upper = upper_;
    md.col_type.resize(n);
    for (std::size_t j = 0; j < n; ++j)
      md.col_type[j] = integer_col_[j] ? VarType::kInteger : VarType::kContinuous;
    // finalize() orders each column by row and drops coefficients under the zero-drop
    // tolerance (1e-11), as every matrix built through SparseMatrix is.
    md.matrix.reset(static_cast<Index>(m), static_cast<Index>(n));
    md.matrix.reserve(entries_.size());
    for (const Entry& e : entries_) md.matrix.add_entry(e.row, e.col, e.value);
    md.matrix.finalize();
    const std::string problem = md.validate();
    if (!problem.empty())
      return ReadResult::failure(format("{}: model failed validation: {}", path_, problem));
    *out = std::move(md);
    return ReadResult::success();
  }

  std::string path_;
  bool fixed_;
  long long line_no_ = 0;
  Section section_ = Section::kNone;
  bool known_header_ = false;
  std::string name_, rhs_vector_, range_vector_, bound_vector_;
  bool maximize_ = false, have_objective_ = false, integer_ = false;
  double objective_rhs_ = 0.0;
  std::unordered_map<std::string, Index> row_index_, col_index_;
  std::unordered_set<Index> objective_seen_;
  std::unordered_set<std::uint64_t> entry_seen_;
  std::vector<char> row_type_, has_range_, lower_explicit_, integer_col_;
  std::vector<std::string> row_names_, col_names_;
  std::vector<double> rhs_, range_, cost_, lower_, upper_;
  std::vector<Entry> entries_;
};

bool ends_with(const std::string& s, const char* suffix) {
  const std::string t(suffix);
  return s.size() >= t.size() &&
         std::equal(t.rbegin(), t.rend(), s.rbegin(), [](char a, char b) {
           return std::tolower(static_cast<unsigned char>(a)) == std::tolower(static_cast<unsigned char>(b));
         });
}

}  // namespace

bool parse_mps_format(const std::string& text, MpsFormat* out) noexcept {
  const std::string t = upper(text);
  if (t != "AUTO" && t != "FREE" && t != "FIXED") return false;
  *out = t == "AUTO" ? MpsFormat::kAuto : t == "FREE" ? MpsFormat::kFree : MpsFormat::kFixed;
  return true;
}

ReadResult read_mps(const std::string& path, Model* model, MpsFormat format_wanted,
                    MpsFormat* format_used) {
  if (model == nullptr) return ReadResult::failure(path + ": no model to read into");
  if (format_wanted != MpsFormat::kAuto) {
    if (format_used != nullptr) *format_used = format_wanted;
    return MpsReader(path, format_wanted).read(model);
  }
  ReadResult free_result = MpsReader(path, MpsFormat::kFree).read(model);
  if (free_result.ok) {
    if (format_used != nullptr) *format_used = MpsFormat::kFree;
    return free_result;
  }
  const ReadResult fixed_result = MpsReader(path, MpsFormat::kFixed).read(model);
  if (fixed_result.ok) {
    if (format_used != nullptr) *format_used = MpsFormat::kFixed;
    return fixed_result;
  }
  return ReadResult::failure(format("{} (the fixed-format reader also failed: {})",
                                    free_result.error, fixed_result.error));
}

ReadResult read_model(const std::string& path, Model* model) {
  std::string p = path;
  if (ends_with(p, ".gz")) p.resize(p.size() - 3);
  if (ends_with(p, ".lp"))
    return ReadResult::failure(path + ": CPLEX LP files are not read by this build; use MPS");
  return read_mps(path, model);  // .mps and any other extension
}

}  // namespace sankhya::io
