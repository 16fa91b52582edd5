// SPDX-License-Identifier: Apache-2.0
// Logger output: leveled lines, the iteration table, the JSONL progress stream (reference
// behaviour: src/util/logging.cpp; JSON written by hand here).
#include "sankhya/logging.hpp"

#include <cmath>

namespace sankhya {

LogLevel log_level_of(const std::string& s) {
  if (s == "off") return LogLevel::kQuiet;
  if (s == "error") return LogLevel::kError;
  if (s == "warning") return LogLevel::kWarning;
  if (s == "verbose") return LogLevel::kVerbose;
  if (s == "debug") return LogLevel::kDebug;
  return LogLevel::kInfo;
}

Logger::~Logger() {
  if (progress_stream_) std::fclose(progress_stream_);
}

void Logger::write(LogLevel, std::string_view line) {
  if (!stream_) return;
  std::fwrite(line.data(), 1, line.size(), stream_);
  std::fputc('\n', stream_);
  std::fflush(stream_);
}

void Logger::begin_iteration_table() {
  rows_since_header_ = 0;
  if (enabled(LogLevel::kInfo))
    write(LogLevel::kInfo, "      Iter        Objective     Primal inf.     Dual inf.     Time");
}

static std::string json_number(double v) {
  if (!std::isfinite(v)) return "null";
  char buf[40];
  std::snprintf(buf, sizeof(buf), "%.17g", v);
  return buf;
}

void Logger::iteration(Count iteration_number, double objective, double primal, double dual,
                       double seconds) {
  if (enabled(LogLevel::kInfo)) {
    if (header_interval_ > 0 && rows_since_header_ >= header_interval_) begin_iteration_table();
    char buf[160];
    std::snprintf(buf, sizeof(buf), "%10lld  %15.8e  %14.6e  %12.6e  %7.2fs",
                  static_cast<long long>(iteration_number), objective, primal, dual, seconds);
    write(LogLevel::kInfo, buf);
    ++rows_since_header_;
  }
  if (progress_stream_) {
    const double t =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - progress_started_).count();
    const std::string line = "{\"event\":\"iteration\",\"phase\":\"lp\",\"iterations\":" +
                             std::to_string(iteration_number) + ",\"objective\":" +
                             json_number(objective) + ",\"best_bound\":" + json_number(objective) +
                             ",\"elapsed\":" + json_number(t) + "}\n";
    std::fwrite(line.data(), 1, line.size(), progress_stream_);
    std::fflush(progress_stream_);
  }
}

void Logger::enable_progress_output(const std::string& path) {
  if (progress_stream_) std::fclose(progress_stream_);
  progress_stream_ = path.empty() ? nullptr : std::fopen(path.c_str(), "a");
  progress_started_ = std::chrono::steady_clock::now();
}

}  // namespace sankhya
