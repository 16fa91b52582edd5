// SPDX-License-Identifier: Apache-2.0
// Logger with the reference's method names (include/sankhya/logging.hpp:40-163): leveled
// lines, the iteration table, and the JSONL progress stream. Patterns use the same "{}" /
// "{:.3e}" placeholders; the engine's formatter (src/text_format.hpp) renders them (libstdc++ 11
// has no <format> and fmt is not a dependency here).
#pragma once

#include <chrono>
#include <cstdio>
#include <string>
#include <string_view>
#include <type_traits>
#include <vector>

#include "sankhya/types.hpp"
#include "../../src/text_format.hpp"

namespace sankhya {

enum class LogLevel { kQuiet = 0, kError = 1, kWarning = 2, kInfo = 3, kVerbose = 4, kDebug = 5 };

/// The `log_level` option's text (off, error, warning, info, verbose, debug) as a level.
[[nodiscard]] LogLevel log_level_of(const std::string& text);

namespace detail {
using hip_engine::text::format;
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
