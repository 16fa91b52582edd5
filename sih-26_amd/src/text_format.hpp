// SPDX-License-Identifier: Apache-2.0
// The engine's own text formatter: renders "{}" and "{:.3e}"-style placeholders (the pattern
// syntax of fmt, which SANKHYA's Logger takes) with snprintf. The engine's status messages use
// it, so the engine sources need nothing beyond SANKHYA's public headers and the standard
// library (libstdc++ 11 has no <format>).
#pragma once

#include <cstdio>
#include <string>
#include <string_view>
#include <type_traits>
#include <vector>

namespace sankhya::hip_engine::text {


inline std::string render_one(const std::string& spec, double v) {
  char buf[64];
  std::string f = "%" + (spec.empty() ? std::string("g") : spec);
  if (f.back() != 'e' && f.back() != 'f' && f.back() != 'g' && f.back() != 'E') f += "g";
  std::snprintf(buf, sizeof(buf), f.c_str(), v);
  return buf;
}
inline std::string render_one(const std::string&, const std::string& v) { return v; }
inline std::string render_one(const std::string&, std::string_view v) { return std::string(v); }
inline std::string render_one(const std::string&, const char* v) { return v ? v : ""; }
inline std::string render_one(const std::string&, bool v) { return v ? "true" : "false"; }
template <typename T, typename std::enable_if<std::is_integral<T>::value && !std::is_same<T, bool>::value, int>::type = 0>
inline std::string render_one(const std::string&, T v) {
  return std::to_string(v);
}
template <typename T, typename std::enable_if<std::is_floating_point<T>::value && !std::is_same<T, double>::value, int>::type = 0>
inline std::string render_one(const std::string& spec, T v) {
  return render_one(spec, static_cast<double>(v));
}

template <typename... Args>
std::string format(std::string_view pattern, const Args&... args) {
  std::vector<std::string> specs;
  std::string out;
  std::vector<std::string> pieces;
  // Split the pattern into literal pieces and placeholder specs.
  std::string lit;
  for (std::size_t i = 0; i < pattern.size(); ++i) {
    const char ch = pattern[i];
    if (ch == '{' && i + 1 < pattern.size() && pattern[i + 1] == '{') {
      lit += '{';
      ++i;
    } else if (ch == '}' && i + 1 < pattern.size() && pattern[i + 1] == '}') {
      lit += '}';
      ++i;
    } else if (ch == '{') {
      const std::size_t close = pattern.find('}', i);
      std::string spec(pattern.substr(i + 1, close - i - 1));
      if (!spec.empty() && spec[0] == ':') spec.erase(0, 1);
      pieces.push_back(lit);
      lit.clear();
      specs.push_back(spec);
      i = close;
    } else {
      lit += ch;
    }
  }
  std::vector<std::string> rendered;
  std::size_t k = 0;
  ((rendered.push_back(render_one(k < specs.size() ? specs[k] : std::string(), args)), ++k), ...);
  for (std::size_t p = 0; p < pieces.size(); ++p) {
    out += pieces[p];
    if (p < rendered.size()) out += rendered[p];
  }
  out += lit;
  return out;
}

}  // namespace sankhya::hip_engine::text
