// SPDX-License-Identifier: Apache-2.0
// Logger with the reference's method names (include/sankhya/logging.hpp:40-163): leveled
// lines, the iteration table, and the JSONL progress stream. Patterns use the same "{}" /
// "{:.3e}" placeholders; a small formatter in this header renders them (libstdc++ 11 has no
// <format> and fmt is not a dependency here).
#pragma once

#include <chrono>
#include <cstdio>
#include <string>
#include <string_view>
#include <type_traits>
#include <vector>

#include "sankhya/types.hpp"

namespace sankhya {

enum class LogLevel { kQuiet = 0, kError = 1, kWarning = 2, kInfo = 3, kVerbose = 4, kDebug = 5 };

/// The `log_level` option's text (off, error, warning, info, verbose, debug) as a level.
[[nodiscard]] LogLevel log_level_of(const std::string& text);

namespace detail {

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

}  // namespace detail

class Profiler;

class Logger {
 public:
  explicit Logger(std::FILE* stream = stdout, LogLevel level = LogLevel::kInfo)
      : stream_(stream), level_(level) {}
  ~Logger();
  Logger(const Logger&) = delete;
  Logger& operator=(const Logger&) = delete;

  void set_level(LogLevel level) noexcept { level_ = level; }
  [[nodiscard]] LogLevel level() const noexcept { return level_; }
  [[nodiscard]] bool enabled(LogLevel level) const noexcept {
    return stream_ != nullptr && static_cast<int>(level) <= static_cast<int>(level_);
  }
  void set_stream(std::FILE* stream) noexcept { stream_ = stream; }
  void write(LogLevel level, std::string_view line);

  template <typename... Args>
  void error(std::string_view p, const Args&... a) { emit(LogLevel::kError, p, a...); }
  template <typename... Args>
  void warning(std::string_view p, const Args&... a) { emit(LogLevel::kWarning, p, a...); }
  template <typename... Args>
  void info(std::string_view p, const Args&... a) { emit(LogLevel::kInfo, p, a...); }
  template <typename... Args>
  void verbose(std::string_view p, const Args&... a) { emit(LogLevel::kVerbose, p, a...); }
  template <typename... Args>
  void debug(std::string_view p, const Args&... a) { emit(LogLevel::kDebug, p, a...); }

  void begin_iteration_table();
  void iteration(Count iteration_number, double objective, double primal_infeasibility,
                 double dual_infeasibility, double seconds);
  void set_iteration_header_interval(int n) noexcept { header_interval_ = n; }
  void enable_progress_output(const std::string& path);

  void set_profiler(Profiler* profiler) noexcept { profiler_ = profiler; }
  [[nodiscard]] Profiler* profiler() const noexcept { return profiler_; }

 private:
  template <typename... Args>
  void emit(LogLevel level, std::string_view p, const Args&... a) {
    if (enabled(level)) write(level, detail::format(p, a...));
  }
  std::FILE* stream_ = nullptr;
  LogLevel level_ = LogLevel::kInfo;
  int header_interval_ = 20;
  int rows_since_header_ = 0;
  std::FILE* progress_stream_ = nullptr;
  Profiler* profiler_ = nullptr;
  std::chrono::steady_clock::time_point progress_started_{};
};

}  // namespace sankhya
