// SPDX-License-Identifier: Apache-2.0
// The solution file (reference: include/sankhya/io.hpp:56-72; the layout of
// src/io/writer.cpp:121-284, restricted to what a PDHG solution carries: no Farkas or ray
// section, no IIS, no solution pool, no ranging).
#include "sankhya/io.hpp"

#include <cmath>
#include <cstdio>
#include <limits>
#include <string>
#include <vector>

#include "sankhya/tolerances.hpp"

namespace sankhya::io {
namespace {

// 17 significant digits read back to the same double; -0 prints as 0 and infinities as words
// (reference writer.cpp:35-39).
std::string number(double v) {
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  if (v == 0.0) v = 0.0;
  char buf[40];
  std::snprintf(buf, sizeof buf, "%.17g", v);
  return buf;
}

// Bare unless it is empty or holds whitespace, a quote or a backslash; then quoted with `"`
// and `\` escaped (reference writer.cpp:76-91).
std::string field_name(const std::string& name) {
  if (!name.empty() && name.find_first_of(" \t\r\n\"\\") == std::string::npos) return name;
  std::string out = "\"";
  for (const char c : name) {
    if (c == '"' || c == '\\') out += '\\';
    out += c;
  }
  out += '"';
  return out;
}

std::string name_or(const std::vector<std::string>& names, Index i, char prefix) {
  const auto u = static_cast<std::size_t>(i);
  if (u < names.size() && !names[u].empty()) return names[u];
  return prefix + std::to_string(i);
}

double at_or_zero(const std::vector<double>& v, Index i) {
  const auto u = static_cast<std::size_t>(i);
  return u < v.size() ? v[u] : 0.0;
}

const char* basis_at(const std::vector<BasisStatus>& v, Index i) {
  const auto u = static_cast<std::size_t>(i);
  return to_string(u < v.size() ? v[u] : BasisStatus::kUnknown);
}

bool write_file(const std::string& path, const Model& model, const Solution& s,
                double mip_relative_gap, double mip_absolute_gap, std::string* error) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (f == nullptr) {
    if (error != nullptr) *error = path + ": cannot open for writing";
    return false;
  }
  const Index n = model.num_cols(), m = model.num_rows();
  auto line = [&](const std::string& key, const std::string& value) {
    std::fprintf(f, "%s %s\n", key.c_str(), value.c_str());
  };
  std::fputs("# SANKHYA solution file\n", f);
  std::fputs("# generated-by SANKHYA restarted-PDHG engine for MI355X (gfx950, HIP)\n", f);
  std::fputs("# All numbers carry 17 significant digits and round-trip exactly.\n", f);
  line("model", model.name.empty() ? "(unnamed)" : model.name);
  line("source", model.source_path);
  line("sense", model.sense == ObjSense::kMaximize ? "maximize" : "minimize");
  line("status", to_string(s.status));
  line("algorithm", s.algorithm.empty() ? "unknown" : s.algorithm);
  line("objective", number(s.objective));
  line("dual_bound", number(s.dual_bound));
  line("mip_relative_gap", number(mip_relative_gap));
  line("mip_absolute_gap", number(mip_absolute_gap));
  line("objective_offset", number(model.objective_offset));
  line("certificate", "none");
  line("rows", std::to_string(m));
  line("columns", std::to_string(n));
  line("iterations", std::to_string(s.iterations));
  line("nodes", "0");
  line("solve_seconds", number(s.solve_seconds));
  line("primal_infeasibility", number(s.primal_infeasibility));
  line("dual_infeasibility", number(s.dual_infeasibility));
  line("integrality_violation", number(s.integrality_violation));
  if (!s.message.empty()) line("message", s.message);

  if (claims_a_point(s.status)) {
    std::fputs("\n# name value reduced_cost basis_status\n", f);
    std::fprintf(f, "begin columns %d\n", static_cast<int>(n));
    for (Index j = 0; j < n; ++j) {
      std::fprintf(f, "%s %s %s %s\n", field_name(name_or(model.col_names, j, 'C')).c_str(),
                   number(at_or_zero(s.col_value, j)).c_str(),
                   number(at_or_zero(s.col_dual, j)).c_str(), basis_at(s.col_status, j));
    }
    std::fputs("end columns\n", f);
    std::fputs("\n# name activity dual basis_status\n", f);
    std::fprintf(f, "begin rows %d\n", static_cast<int>(m));
    for (Index i = 0; i < m; ++i) {
      std::fprintf(f, "%s %s %s %s\n", field_name(name_or(model.row_names, i, 'R')).c_str(),
                   number(at_or_zero(s.row_activity, i)).c_str(),
                   number(at_or_zero(s.row_dual, i)).c_str(), basis_at(s.row_status, i));
    }
    std::fputs("end rows\n", f);
  }
  const bool ok = std::ferror(f) == 0;
  const bool closed = std::fclose(f) == 0;
  if (!(ok && closed)) {
    if (error != nullptr) *error = path + ": write failed";
    return false;
  }
  return true;
}

}  // namespace

bool write_solution(const std::string& path, const Model& model, const Solution& solution,
                    std::string* error) {
  return write_file(path, model, solution, tol::kMipRelativeGap, tol::kMipAbsoluteGap, error);
}

bool write_solution(const std::string& path, const Model& model, const Solution& solution,
                    const Options& options, std::string* error) {
  return write_file(path, model, solution, options.get_double("mip_relative_gap"),
                    options.get_double("mip_absolute_gap"), error);
}

}  // namespace sankhya::io
